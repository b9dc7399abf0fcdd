// rmq_device.hpp — device helpers for the gfx950 RMQ kernels.
//
// Wave64 everywhere: group reductions use DPP row operations (quad_perm, row_half_mirror,
// row_mirror stay inside 4/8/16-lane groups of a 16-lane row), one ds_swizzle for the
// 32-lane step and one ds_bpermute for the 64-lane step. No 32-lane warp assumptions.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace rmq {

constexpr uint32_t kMaxLevels = 20;  // XXX_MAX_LEVELS, src/constants.cuh:14

// Level metadata as seen by the kernels (the reference's xxx_metadata_type,
// src/constants.cuh:34-39, plus the level-buffer pointers). Passed by value.
struct HierView {
  const float* A;        // level 0: the caller's array (borrowed)
  const float* lv;       // level buffer, values  (levels 1..num_levels-1)
  const uint8_t* lr;     // winner offsets of levels 1.. (rel_load), at byte offsets roff[k]
  const uint32_t* zeros; // which signs of zero level 0 holds (kZeroPos | kZeroNeg), see below
  const float* tail;     // level 0's last 16-byte vector A[n & ~3, n), +inf padded, 16-byte aligned
  const uint2* tt;       // sparse table over the top level (dev::top_lookup), or null
  // block scans of levels ps_lo.., or null: entry i of level k at ps[off[k] + i] holds {value
  // bits, level-0 index} of the leftmost minimum of the rest of its R-block [i, block end) in
  // .x/.y and of the block up to it [block start, i] in .z/.w
  const uint4* ps;
  uint32_t ps_lo;        // the lowest level with block scans
  uint32_t tstride;      // entries per row of tt
  uint32_t num_levels;
  uint32_t size[kMaxLevels];
  uint32_t off[kMaxLevels];
  uint32_t roff[kMaxLevels];
};

// Where the leftmost minimum of an upper-level entry lies: entry j of level k >= 1 keeps
// i - (j << (R_log k)), i = the leftmost level-0 index of the block's minimum, in the fewest
// whole bytes that hold a block's span (R^k <= 2^8: one byte, <= 2^16: two, else four). A level's
// offsets start 16-byte aligned and are padded by four entries.
__host__ __device__ constexpr uint32_t rel_bytes(uint32_t span_log) {
  return span_log <= 8 ? 1u : (span_log <= 16 ? 2u : 4u);
}

// A value-only answer must carry the bit pattern of A[leftmost argmin]; a plain minimum can only
// miss it when the minimum is zero and the array holds both +0 and -0. The level-0 build pass
// records which signs occur, and whether the array holds a NaN (kHasNaN: queries then apply
// CPU BASE's NaN rule, dev::nan_at_l); a hierarchy with no such pass (one level) records all.
constexpr uint32_t kZeroPos = 1u, kZeroNeg = 2u, kHasNaN = 4u;
constexpr uint32_t kLevel0Unknown = kZeroPos | kZeroNeg | kHasNaN;

namespace dev {

constexpr uint32_t kNoIdx = 0xFFFFFFFFu;
#define RMQ_INF __builtin_huge_valf()

__device__ __forceinline__ uint32_t f2u(float x) { return __builtin_bit_cast(uint32_t, x); }
__device__ __forceinline__ float u2f(uint32_t x) { return __builtin_bit_cast(float, x); }

// DPP control codes (GCN3+ encoding): quad_perm = p0 | p1<<2 | p2<<4 | p3<<6.
enum : int {
  kDppXor1 = 0xB1,         // quad_perm [1,0,3,2]
  kDppXor2 = 0x4E,         // quad_perm [2,3,0,1]
  kDppHalfMirror = 0x141,  // lane i <- lane 7-i inside each 8-lane half row
  kDppMirror = 0x140,      // lane i <- lane 15-i inside each 16-lane row
};

template <int CTRL>
__device__ __forceinline__ uint32_t dpp_mov(uint32_t x) {
  return static_cast<uint32_t>(
      __builtin_amdgcn_update_dpp(0, static_cast<int>(x), CTRL, 0xF, 0xF, false));
}

// Partner value for step `S` of a W-lane all-reduce (S = 1, 2, 4, 8, 16, 32 = xor distance,
// realised with mirrors, which pair the same halves as xor once the smaller groups agree).
template <int S>
__device__ __forceinline__ uint32_t partner(uint32_t x) {
  if constexpr (S == 1) return dpp_mov<kDppXor1>(x);
  if constexpr (S == 2) return dpp_mov<kDppXor2>(x);
  if constexpr (S == 4) return dpp_mov<kDppHalfMirror>(x);
  if constexpr (S == 8) return dpp_mov<kDppMirror>(x);
  if constexpr (S == 16)  // ds_swizzle bit mode: and 0x1F, or 0, xor 0x10
    return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(x), 0x401F));
  if constexpr (S == 32) return static_cast<uint32_t>(__shfl_xor(static_cast<int>(x), 32));
  return x;
}

// Candidate order: smaller value, then smaller index. Scanning left to right with a strict
// `<` (CPU BASE, src/cpu_methods.h:43-52) selects exactly the minimum of this order.
__device__ __forceinline__ bool better(float v, uint32_t i, float bv, uint32_t bi) {
  return (v < bv) || (v == bv && i < bi);
}

__device__ __forceinline__ float fmin_sel(float a, float b) { return b < a ? b : a; }

template <int S>
__device__ __forceinline__ void step_min(float& v) {
  v = fmin_sel(v, u2f(partner<S>(f2u(v))));
}

template <int S>
__device__ __forceinline__ void step_argmin(float& v, uint32_t& i) {
  const float ov = u2f(partner<S>(f2u(v)));
  const uint32_t oi = partner<S>(i);
  const bool take = better(ov, oi, v, i);
  v = take ? ov : v;
  i = take ? oi : i;
}

// All-reduce over aligned groups of W lanes; every lane of the group gets the result.
// All W lanes of a group must be active (callers keep control flow group-uniform).
template <int W>
__device__ __forceinline__ float group_min(float v) {
  if constexpr (W >= 2) step_min<1>(v);
  if constexpr (W >= 4) step_min<2>(v);
  if constexpr (W >= 8) step_min<4>(v);
  if constexpr (W >= 16) step_min<8>(v);
  if constexpr (W >= 32) step_min<16>(v);
  if constexpr (W >= 64) step_min<32>(v);
  return v;
}

template <int W>
__device__ __forceinline__ void group_argmin(float& v, uint32_t& i) {
  if constexpr (W >= 2) step_argmin<1>(v, i);
  if constexpr (W >= 4) step_argmin<2>(v, i);
  if constexpr (W >= 8) step_argmin<4>(v, i);
  if constexpr (W >= 16) step_argmin<8>(v, i);
  if constexpr (W >= 32) step_argmin<16>(v, i);
  if constexpr (W >= 64) step_argmin<32>(v, i);
}

// 16-byte load of base[i..i+4) where i % 4 == 0 and i < size; elements at or past `size`
// read as +inf and are never touched in memory (no over-read of the caller's array).
__device__ __forceinline__ float4 load4(const float* __restrict__ base, uint32_t i, uint32_t size) {
  if (i + 4 <= size) return *reinterpret_cast<const float4*>(base + i);
  float4 r = make_float4(RMQ_INF, RMQ_INF, RMQ_INF, RMQ_INF);
  r.x = base[i];
  if (i + 1 < size) r.y = base[i + 1];
  if (i + 2 < size) r.z = base[i + 2];
  return r;
}

// Streaming form for data read once per query (level-0 blocks, the query pairs): the
// non-temporal hint keeps it from displacing the upper levels in L2.
#ifndef RMQ_STREAM_L0
#define RMQ_STREAM_L0 1
#endif
typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef uint32_t u32x2 __attribute__((ext_vector_type(2)));

__device__ __forceinline__ float4 load4_stream(const float* __restrict__ base, uint32_t i,
                                               uint32_t size) {
#if RMQ_STREAM_L0
  if (i + 4 <= size) {
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(base + i));
    return make_float4(v.x, v.y, v.z, v.w);
  }
#endif
  return load4(base, i, size);
}

__device__ __forceinline__ uint2 load_query(const uint2* __restrict__ Q, uint64_t qi) {
#if RMQ_STREAM_L0
  const u32x2 v = __builtin_nontemporal_load(reinterpret_cast<const u32x2*>(Q + qi));
  return make_uint2(v.x, v.y);
#else
  return Q[qi];
#endif
}

// Element j (0..3) of a vector.
__device__ __forceinline__ float pick4(float4 x, uint32_t j) {  // two selects, no branch
  const float lo = (j & 1u) ? x.y : x.x;
  const float hi = (j & 1u) ? x.w : x.z;
  return (j & 2u) ? hi : lo;
}

// The copy of level 0's last vector kept with a hierarchy: tail[j] = A[(n & ~3) + j] below n,
// +inf past it (one lane writes all four).
__device__ __forceinline__ void fill_tail(const float* __restrict__ A, uint32_t n, float* __restrict__ tail) {
  const uint32_t p = n & ~3u;
#pragma unroll
  for (uint32_t j = 0; j < 4; ++j) tail[j] = p + j < n ? A[p + j] : RMQ_INF;
}

// What x tells about level 0, if `ok`: the sign bit of a zero (kZeroPos / kZeroNeg), kHasNaN
// for a NaN.
__device__ __forceinline__ uint32_t zero_bits(float x, bool ok) {
  const uint32_t b = __float_as_uint(x);
  const uint32_t z = (b & 0x7FFFFFFFu) == 0 ? (1u << (b >> 31)) : 0u;
  return ok ? (z | (x != x ? kHasNaN : 0u)) : 0u;
}

// A zero minimum v of a value-only query, given the array's zero signs: true (and v set to the
// right zero) unless both signs occur, when the query must be answered with index tracking.
__device__ __forceinline__ bool zero_known(uint32_t zeros, float& v) {
  if ((zeros & (kZeroPos | kZeroNeg)) == (kZeroPos | kZeroNeg)) return false;
  v = (zeros & kZeroNeg) ? -0.0f : 0.0f;
  return true;
}

// OR of the wave's `bits` (kZeroPos | kZeroNeg | kHasNaN) into *zeros, one atomic per wave that
// saw any. All lanes of the wave must call it.
__device__ __forceinline__ void publish_zeros(uint32_t* zeros, uint32_t bits) {
  const uint64_t pos = __ballot((bits & kZeroPos) != 0u);
  const uint64_t neg = __ballot((bits & kZeroNeg) != 0u);
  const uint64_t nan = __ballot((bits & kHasNaN) != 0u);
  const uint32_t w = (pos ? kZeroPos : 0u) | (neg ? kZeroNeg : 0u) | (nan ? kHasNaN : 0u);
  if (w && (threadIdx.x & 63u) == 0) atomicOr(zeros, w);
}

// CPU BASE's NaN rule (src/cpu_methods.h:7-16, 42-53): the scan starts from A[l] and only a
// strictly smaller element replaces it, so a NaN at A[l] is the answer and a NaN anywhere else
// never is. The solvers never take a NaN (every comparison with one is false), so they give the
// answer in every other case; this supplies the first, on arrays whose level-0 pass saw a NaN
// (flags: the hierarchy's level-0 word) and only there loads A[l]. True (and v = A[l]) when it
// applies.
__device__ __forceinline__ bool nan_at_l(uint32_t flags, const float* __restrict__ A, uint64_t l, float& v) {
  if (!(flags & kHasNaN)) return false;
  const float x = A[l];
  if (x == x) return false;
  v = x;
  return true;
}

// Final ranges on the top level longer than this are answered from the sparse table.
constexpr uint32_t kTopScan = 8;

// Minimum of top-level entries [a, b), b - a >= 2, from the sparse table: the entry i of row
// j >= 1 (at tt + (j - 1) * stride) holds {value bits, leftmost position} of the minimum of
// [i, i + 2^j). With j = floor(log2(b - a)) the entries at a and b - 2^j cover [a, b); on equal
// values the left one holds the leftmost position, since any position of that value in the
// right entry's span left of it lies in the left entry's span too.
// Entry b replaces entry a (a left of b) in a sparse-table row: strictly smaller, or a holds a NaN
// and b does not (a NaN never wins; ties keep the left entry, which holds the leftmost position).
__device__ __forceinline__ bool top_take_right(float a, float b) { return b < a || (a != a && b == b); }

// The minimum of top-level entries [a, b) (b - a >= 2) from two table entries of row
// floor(log2(b - a)): its value and the entry's payload (the level-0 index of the leftmost
// minimum); on equal values the left entry holds the leftmost one.
__device__ __forceinline__ void top_lookup(const uint2* __restrict__ tt, uint32_t stride, uint32_t a,
                                           uint32_t b, float& v, uint32_t& pos) {
  const uint32_t j = 31u - static_cast<uint32_t>(__clz(b - a));
  const uint2* row = tt + static_cast<size_t>(j - 1u) * stride;
  const uint2 x = row[a];
  const uint2 y = row[b - (1u << j)];
  const bool right = u2f(y.x) < u2f(x.x);
  v = u2f(right ? y.x : x.x);
  pos = right ? y.y : x.y;
}

__device__ __forceinline__ bool in_range(uint32_t p, uint32_t lo, uint32_t hi) {
  return p - lo < hi - lo;  // lo <= p < hi, with lo <= hi
}

// Fold the elements of a float4 at position i that fall in [lo, hi) into a running minimum.
__device__ __forceinline__ float fold4(float acc, float4 x, uint32_t i, uint32_t lo, uint32_t hi) {
  acc = in_range(i + 0, lo, hi) ? fmin_sel(acc, x.x) : acc;
  acc = in_range(i + 1, lo, hi) ? fmin_sel(acc, x.y) : acc;
  acc = in_range(i + 2, lo, hi) ? fmin_sel(acc, x.z) : acc;
  acc = in_range(i + 3, lo, hi) ? fmin_sel(acc, x.w) : acc;
  return acc;
}

__device__ __forceinline__ void take(float& bv, uint32_t& bi, float v, uint32_t i, bool ok) {
  const bool t = ok && better(v, i, bv, bi);
  bv = t ? v : bv;
  bi = t ? i : bi;
}

// ---- index mode without index-array traffic ----------------------------------------------
// The pieces a query visits stand for disjoint level-0 spans, so candidates are totally ordered
// by the level-0 start of their span ("key" = position << (R_log * level)). Comparing
// (value, key) therefore realises the leftmost rule without reading any stored index; only the
// winner's index is looked up at the end. "pay" says where that index lives: kDirect for a
// level-0 candidate (its key is its position), else the level k of the candidate.
constexpr uint32_t kDirect = 0x80000000u;

struct Cand {
  float v;
  uint32_t key;
  uint32_t pay;
};

__device__ __forceinline__ Cand cand_none() { return Cand{RMQ_INF, kNoIdx, kNoIdx}; }

__device__ __forceinline__ void cand_take(Cand& b, float v, uint32_t key, uint32_t pay, bool ok) {
  const bool t = ok && better(v, key, b.v, b.key);
  b.v = t ? v : b.v;
  b.key = t ? key : b.key;
  b.pay = t ? pay : b.pay;
}

// Fold the elements of the vector at level position p that fall in [lo, hi). shift =
// R_log * level; pay = kDirect at level 0, else the level.
__device__ __forceinline__ void cand_fold4(Cand& b, float4 x, uint32_t p, uint32_t lo, uint32_t hi,
                                           uint32_t shift, uint32_t pay) {
  cand_take(b, x.x, (p + 0) << shift, pay, in_range(p + 0, lo, hi));
  cand_take(b, x.y, (p + 1) << shift, pay, in_range(p + 1, lo, hi));
  cand_take(b, x.z, (p + 2) << shift, pay, in_range(p + 2, lo, hi));
  cand_take(b, x.w, (p + 3) << shift, pay, in_range(p + 3, lo, hi));
}

template <int S>
__device__ __forceinline__ void step_cand(Cand& b) {
  const float ov = u2f(partner<S>(f2u(b.v)));
  const uint32_t ok = partner<S>(b.key);
  const uint32_t op = partner<S>(b.pay);
  const bool t = better(ov, ok, b.v, b.key);
  b.v = t ? ov : b.v;
  b.key = t ? ok : b.key;
  b.pay = t ? op : b.pay;
}

// The value of the first lane of each aligned group of W lanes (all W lanes active), in every
// lane of the group: a quad permute (W = 4), a ds_swizzle and-mask (W = 8 .. 32), or the
// wavefront's first lane (W = 64).
template <int W>
__device__ __forceinline__ uint32_t group_first(uint32_t x) {
  static_assert(W == 4 || W == 8 || W == 16 || W == 32 || W == 64, "group of 4 .. 64 lanes");
  if constexpr (W == 4) return dpp_mov<0x00>(x);  // quad_perm [0,0,0,0]
  else if constexpr (W == 64) return static_cast<uint32_t>(__builtin_amdgcn_readfirstlane(static_cast<int>(x)));
  else return static_cast<uint32_t>(__builtin_amdgcn_ds_swizzle(static_cast<int>(x), 0x1F & ~(W - 1)));
}

// A group's candidate when only its first lane can hold a better one than the others (which
// then all hold the same): that lane's, in every lane, without the reduction.
template <int W>
__device__ __forceinline__ void group_first_cand(Cand& b) {
  b.v = u2f(group_first<W>(f2u(b.v)));
  b.key = group_first<W>(b.key);
  b.pay = group_first<W>(b.pay);
}

template <int W>
__device__ __forceinline__ void group_cand(Cand& b) {
  if constexpr (W >= 2) step_cand<1>(b);
  if constexpr (W >= 4) step_cand<2>(b);
  if constexpr (W >= 8) step_cand<4>(b);
  if constexpr (W >= 16) step_cand<8>(b);
  if constexpr (W >= 32) step_cand<16>(b);
  if constexpr (W >= 64) step_cand<32>(b);
}

// Offset entry j of a level whose blocks span 2^span_log level-0 elements (roff: the level's
// byte offset in lr). One aligned 4-byte load; the entry is cut out of it.
__device__ __forceinline__ uint32_t rel_load(const uint8_t* __restrict__ lr, uint32_t roff, uint32_t j,
                                             uint32_t span_log) {
  const uint32_t w = rel_bytes(span_log);
  const uint32_t a = roff + j * w;
  const uint32_t word = *reinterpret_cast<const uint32_t*>(lr + (a & ~3u));
  const uint32_t mask = w == 4u ? 0xFFFFFFFFu : (1u << (8u * w)) - 1u;
  return (word >> (8u * (a & 3u))) & mask;
}

// Four consecutive offset entries (pos % 4 == 0) of a level, one 4-, 8- or 16-byte load; the
// level's padding covers the entries past its end.
__device__ __forceinline__ uint4 rel_load4(const uint8_t* __restrict__ r, uint32_t pos, uint32_t w) {
  if (w == 1u) {
    const uint32_t x = *reinterpret_cast<const uint32_t*>(r + pos);
    return make_uint4(x & 0xFFu, (x >> 8) & 0xFFu, (x >> 16) & 0xFFu, x >> 24);
  }
  if (w == 2u) {
    const uint2 x = *reinterpret_cast<const uint2*>(r + 2u * pos);
    return make_uint4(x.x & 0xFFFFu, x.x >> 16, x.y & 0xFFFFu, x.y >> 16);
  }
  return *reinterpret_cast<const uint4*>(r + 4u * pos);
}

// Stores the offset entry of level position j (blocks of 2^span_log elements) for winner bi.
__device__ __forceinline__ void rel_store(uint8_t* __restrict__ drel, uint32_t j, uint32_t span_log, uint32_t bi) {
  const uint32_t w = rel_bytes(span_log);
  const uint32_t rel = bi - (span_log >= 32u ? 0u : (j << span_log));
  if (w == 1u) drel[j] = static_cast<uint8_t>(rel);
  else if (w == 2u) reinterpret_cast<uint16_t*>(drel)[j] = static_cast<uint16_t>(rel);
  else reinterpret_cast<uint32_t*>(drel)[j] = rel;
}

}  // namespace dev
}  // namespace rmq
