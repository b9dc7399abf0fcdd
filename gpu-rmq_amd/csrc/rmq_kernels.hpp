// rmq_kernels.hpp — gfx950 kernels of the hierarchical RMQ: level build and the three query
// variants (algs 16, 24, 20).
//
// Data layout in HBM: level 0 is the caller's float array (borrowed). Levels 1.. live in one
// value buffer at the reference's offsets (multiples of 32 elements, src/xxx.h:238), so every
// level starts 128-byte aligned and every R-block of R >= 4 elements is whole 16-byte vectors;
// where each entry's minimum lies is kept beside them as a 1-, 2- or 4-byte offset in its block
// (rel_bytes in rmq_device.hpp).
#pragma once

#include "rmq_device.hpp"

namespace rmq {

constexpr int kBlock = 256;  // 4 waves per workgroup

// ---------------------------------------------------------------------------------------
// Build: level k+1 entry j = min of level k over [j*R, min(j*R+R, size_k)) with the leftmost
// level-0 index of that minimum (reference: xxx_build_step, src/xxx.h:91-120, one launch per
// level, which keeps the value only), stored as an offset entry (rel_store).
//
// Here several levels are made per pass. One workgroup reduces a tile of 2^SLOG = 4096 level-k
// entries (whole blocks of level k + M): sweeps of 1024 entries, each lane a 16-byte vector (one wave
// instruction streams 1 KiB of contiguous level-k data), W = R/4 lanes meeting in the DPP
// argmin, give the tile's level-(k+1) entries, which are written out and kept in LDS; levels
// k+2 .. k+m are then reduced in LDS, one thread per entry scanning its R sources left to
// right, and written as they are made. Level k is read once for m levels, where one pass per
// level would read every level once more for the next.
constexpr uint32_t kPassTileLog = 12;  // level-k entries per build tile (log2): four sweeps

template <int RL>
struct PassShape {
  static constexpr uint32_t SLOG = kPassTileLog;
  static constexpr uint32_t M = SLOG / RL;    // levels per pass: 6 at R = 4, 2 at R = 32
  static constexpr uint32_t LDS = 1u << (SLOG - RL);  // level-(k+1) entries of one tile
  // the LDS levels give one thread each entry of level k+2 and up (tiles of 2^13 entries at R = 4
  // would leave half of a tile's level-(k+2) entries unmade)
  static_assert(M < 2 || LDS / (1u << RL) <= kBlock, "level k+2 of a tile must fit one workgroup");
};

// Tile `tile` of the pass from level k of H making levels k+1 .. k+m (m <= M). s_v / s_i hold
// PassShape<RL>::LDS entries. Level 0 (k = 0) also ORs its zero-sign / NaN bits into *H.zeros.
template <int RL>
__device__ __forceinline__ void build_pass_tile(const HierView& H, uint32_t k, uint32_t m, uint32_t tile,
                                                float* s_v, uint32_t* s_i) {
  using P = PassShape<RL>;
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t W = R / 4;                        // lanes per block of R entries (<= 32)
  constexpr uint32_t kSweep = kBlock * 4;              // level-k entries per sweep
  constexpr uint32_t kSweeps = (1u << P::SLOG) / kSweep;
  constexpr uint32_t kPerSweep = kSweep / R;           // level-(k+1) entries per sweep
  const bool l0 = k == 0;
  const float* sv = l0 ? H.A : H.lv + H.off[k];
  const uint8_t* srel = l0 ? nullptr : H.lr + H.roff[k];
  const uint32_t ssize = H.size[k];
  const uint32_t t0 = tile << P::SLOG;
  float* dv = const_cast<float*>(H.lv) + H.off[k + 1];
  uint8_t* drel = const_cast<uint8_t*>(H.lr) + H.roff[k + 1];
  const uint32_t dsize = H.size[k + 1];
  const uint32_t gl = threadIdx.x & (W - 1);
  const uint32_t sw = rel_bytes(RL * k);
  uint32_t zb = 0;
  // kBatch sweeps' loads are issued before any of them is reduced
  constexpr uint32_t kBatch = kSweeps < 4 ? kSweeps : 4;
  for (uint32_t w0 = 0; w0 < kSweeps; w0 += kBatch) {
    float4 x[kBatch];
    uint4 id[kBatch];
#pragma unroll
    for (uint32_t b = 0; b < kBatch; ++b) {
      const uint32_t pos = t0 + (w0 + b) * kSweep + 4 * threadIdx.x;
      x[b] = make_float4(RMQ_INF, RMQ_INF, RMQ_INF, RMQ_INF);
      id[b] = make_uint4(0, 0, 0, 0);
      if (pos < ssize) {
        x[b] = dev::load4(sv, pos, ssize);
        if (!l0) id[b] = dev::rel_load4(srel, pos, sw);
      }
    }
#pragma unroll
    for (uint32_t b = 0; b < kBatch; ++b) {
      const uint32_t pos = t0 + (w0 + b) * kSweep + 4 * threadIdx.x;
      const uint32_t sh = RL * k;  // level-0 span of a level-k entry (log2)
      float bv = RMQ_INF;
      uint32_t bi = dev::kNoIdx;
      dev::take(bv, bi, x[b].x, id[b].x + ((pos + 0) << sh), pos + 0 < ssize);
      dev::take(bv, bi, x[b].y, id[b].y + ((pos + 1) << sh), pos + 1 < ssize);
      dev::take(bv, bi, x[b].z, id[b].z + ((pos + 2) << sh), pos + 2 < ssize);
      dev::take(bv, bi, x[b].w, id[b].w + ((pos + 3) << sh), pos + 3 < ssize);
      if (l0)
        zb |= dev::zero_bits(x[b].x, pos < ssize) | dev::zero_bits(x[b].y, pos + 1 < ssize) |
              dev::zero_bits(x[b].z, pos + 2 < ssize) | dev::zero_bits(x[b].w, pos + 3 < ssize);
      dev::group_argmin<W>(bv, bi);
      const uint32_t lj = (w0 + b) * kPerSweep + threadIdx.x / W;
      const uint32_t j = (t0 >> RL) + lj;
      if (gl == 0) {
        if (j < dsize) {
          // a level-0 block of NaNs (no element is ever taken): its first position, value +inf,
          // which never wins over a real element left of it and is never asked for on its own
          // (a query whose range holds no other element starts at a NaN, dev::nan_at_l)
          if (l0 && bi == dev::kNoIdx) bi = j * R;
          dv[j] = bv;
          dev::rel_store(drel, j, RL * (k + 1), bi);
        }
        s_v[lj] = bv;
        s_i[lj] = bi;
      }
    }
  }
  if (l0) dev::publish_zeros(const_cast<uint32_t*>(H.zeros), zb);
  uint32_t cnt = P::LDS;
  for (uint32_t L = 2; L <= m; ++L) {
    __syncthreads();
    cnt >>= RL;
    float v = RMQ_INF;
    uint32_t i = dev::kNoIdx;
    if (threadIdx.x < cnt) {
#pragma unroll 8
      for (uint32_t r = 0; r < R; ++r) dev::take(v, i, s_v[threadIdx.x * R + r], s_i[threadIdx.x * R + r], true);
    }
    __syncthreads();
    if (threadIdx.x < cnt) {
      s_v[threadIdx.x] = v;
      s_i[threadIdx.x] = i;
      const uint32_t j = (tile << (P::SLOG - RL * L)) + threadIdx.x;
      if (j < H.size[k + L]) {
        const_cast<float*>(H.lv)[H.off[k + L] + j] = v;
        dev::rel_store(const_cast<uint8_t*>(H.lr) + H.roff[k + L], j, RL * (k + L), i);
      }
    }
  }
}

// One pass of a plain build: tiles 0 .. gridDim.x - 1 of level k of H.
template <int RL>
__global__ __launch_bounds__(kBlock) void k_build_pass(HierView H, uint32_t k, uint32_t m) {
  __shared__ float s_v[PassShape<RL>::LDS];
  __shared__ uint32_t s_i[PassShape<RL>::LDS];
  build_pass_tile<RL>(H, k, m, blockIdx.x, s_v, s_i);
}

// The sparse table over a top level top[0, m), m <= kTopSmall, built by one workgroup of
// kTopThreads threads: row j >= 1 entry i (i + 2^j <= m) is the leftmost minimum of
// [i, i + 2^j) (dev::top_lookup). The current row lives in the LDS array `cur` and is rebuilt in
// place, each thread reading its entries' two sources before a barrier and writing after it.
constexpr uint32_t kTopSmall = 8192;
constexpr uint32_t kTopThreads = 1024;

// The level-0 index of the minimum under entry i of a level whose blocks span 2^sk level-0
// elements (its offset entries at lr + roff); entry i itself when the level is level 0 (lr null).
__device__ __forceinline__ uint32_t level0_index(const uint8_t* __restrict__ lr, uint32_t roff, uint32_t i,
                                                 uint32_t sk) {
  return lr ? (i << sk) + dev::rel_load(lr, roff, i, sk) : i;
}

// Table entries are {value bits, level-0 index of the leftmost minimum}, so a query answered from
// the table needs no offset entry afterwards (top: the level's values; lr / roff / sk: its
// offset entries for level0_index).
__device__ __forceinline__ void top_table_block(const float* __restrict__ top, uint32_t m, uint2* __restrict__ tt,
                                                uint32_t stride, uint32_t rows, uint2* cur,
                                                const uint8_t* __restrict__ lr, uint32_t roff, uint32_t sk) {
  constexpr uint32_t kPer = kTopSmall / kTopThreads;
  for (uint32_t i = threadIdx.x; i < m; i += kTopThreads)
    cur[i] = make_uint2(dev::f2u(top[i]), level0_index(lr, roff, i, sk));
  __syncthreads();
  for (uint32_t j = 1; j <= rows; ++j) {
    const uint32_t h = 1u << (j - 1u);
    const uint32_t cnt = m + 1u - 2u * h;  // entries i with i + 2^j <= m
    uint2 nv[kPer];
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t i = threadIdx.x + r * kTopThreads;
      if (i < cnt) {
        const uint2 a = cur[i], b = cur[i + h];
        nv[r] = dev::top_take_right(dev::u2f(a.x), dev::u2f(b.x)) ? b : a;
      }
    }
    __syncthreads();
    uint2* row = tt + static_cast<size_t>(j - 1u) * stride;
#pragma unroll
    for (uint32_t r = 0; r < kPer; ++r) {
      const uint32_t i = threadIdx.x + r * kTopThreads;
      if (i < cnt) {
        cur[i] = nv[r];
        row[i] = nv[r];
      }
    }
    __syncthreads();
  }
}

// Block scans of level k >= 1 (RMQ_BUILD_BLOCK_SCANS): thread b takes R-block b of level k,
// entries [bR, min(bR + R, size_k)), and writes for each entry i the leftmost minimum of
// [i, block end) (right to left, a left entry wins ties) and of [block start, i] (left to
// right, strict <), each as {value bits, level-0 index}. Upper levels hold no NaN.
template <int RL>
__global__ __launch_bounds__(kBlock) void k_block_scans(HierView H, uint32_t k, uint4* __restrict__ ps) {
  constexpr uint32_t R = 1u << RL;
  const uint32_t size = H.size[k];
  const uint32_t b = blockIdx.x * kBlock + threadIdx.x;
  if (b >= (size + R - 1u) >> RL) return;
  const uint32_t base = b << RL;
  const uint32_t cnt = min(R, size - base);
  const float* __restrict__ v = H.lv + H.off[k];
  const uint32_t sk = RL * k;
  uint2* __restrict__ out = reinterpret_cast<uint2*>(ps + H.off[k] + base);
  float bv = RMQ_INF;
  uint32_t bi = 0;
  for (uint32_t j = 0; j < cnt; ++j) {
    const float x = v[base + j];
    if (j == 0 || x < bv) {
      bv = x;
      bi = level0_index(H.lr, H.roff[k], base + j, sk);
    }
    out[2u * j + 1u] = make_uint2(dev::f2u(bv), bi);
  }
  for (uint32_t j = cnt; j-- > 0;) {
    const float x = v[base + j];
    if (j == cnt - 1u || x <= bv) {
      bv = x;
      bi = level0_index(H.lr, H.roff[k], base + j, sk);
    }
    out[2u * j] = make_uint2(dev::f2u(bv), bi);
  }
}

// The same scans with one lane per entry (R <= 64: a block is an aligned segment of R lanes of a
// wavefront): log-step scans across the segment, leftwards for the prefix (a left partner wins
// ties) and rightwards for the suffix (a right partner must be strictly smaller), then one
// 16-byte store per lane, so a wavefront writes 1 KiB of consecutive entries. Lanes past the
// level hold +inf and store nothing; no entry of the level reads them (a prefix looks left,
// and a right partner of +inf never replaces an entry).
template <int RL>
__global__ __launch_bounds__(kBlock) void k_block_scans_lanes(HierView H, uint32_t k, uint4* __restrict__ ps) {
  constexpr int R = 1 << RL;
  static_assert(R <= 64, "a block within one wavefront");
  const uint32_t size = H.size[k];
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t j = i & static_cast<uint32_t>(R - 1);
  const bool live = i < size;
  const float x = live ? H.lv[H.off[k] + i] : RMQ_INF;
  const uint32_t xi = live ? level0_index(H.lr, H.roff[k], i, RL * k) : 0u;
  float pv = x, sv = x;
  uint32_t pi = xi, si = xi;
#pragma unroll
  for (int d = 1; d < R; d <<= 1) {
    const float lv = dev::u2f(static_cast<uint32_t>(__shfl_up(static_cast<int>(dev::f2u(pv)), d, R)));
    const uint32_t li = static_cast<uint32_t>(__shfl_up(static_cast<int>(pi), d, R));
    const float rv = dev::u2f(static_cast<uint32_t>(__shfl_down(static_cast<int>(dev::f2u(sv)), d, R)));
    const uint32_t ri = static_cast<uint32_t>(__shfl_down(static_cast<int>(si), d, R));
    if (j >= static_cast<uint32_t>(d) && lv <= pv) {
      pv = lv;
      pi = li;
    }
    if (j + static_cast<uint32_t>(d) < static_cast<uint32_t>(R) && rv < sv) {
      sv = rv;
      si = ri;
    }
  }
  if (live) ps[H.off[k] + i] = make_uint4(dev::f2u(sv), si, dev::f2u(pv), pi);
}

// ---------------------------------------------------------------------------------------
// Query traversal shared by algs 16 and 24 (src/xxx.h:158-187 and 493-521): r exclusive;
// while the remaining range at this level is longer than 2R, the partial blocks
// [l, ceil_R(l)) and [floor_R(r), r) belong to the answer and the walk climbs; the range
// left at the level reached (the "final range") is scanned. A group of G lanes works on one
// query: each lane owns 16-byte vectors, so a 32-float block is one vector load per lane at
// G = 8.
//
// Two evaluation orders of the same pieces, same answers:
//  - solve_cl (full): every piece is read, level by level, as the reference does.
//  - solve_cl_pruned (default): the final range is read first. All left pieces lie inside the
//    level-K block just left of the final range, all right pieces inside the block just right
//    of it ("guards", adjacent to the final range, so usually on the same cache lines). A tail
//    is read only if its guard could still win: left tail if guard <= best (a left element
//    wins ties), right tail if guard < best. Walking down, the next guard is the level-k entry
//    just outside the piece just read. For random data the tails are rarely needed, so most
//    queries touch only the small top of the hierarchy instead of two random level-0 lines.
//
// Closed forms of the reference's climb: l_k = ceil(l / R^k), e_k = floor(e / R^k).

// Index of the winner / value-only result of a group, written by all lanes of the group.
struct GroupAnswer {
  float v;
  uint32_t idx;
};

__device__ __forceinline__ const float* level_values(const HierView& H, uint32_t k) {
  return k == 0 ? H.A : H.lv + H.off[k];
}

// Where a solver finds a level's offset and size (one lookup per level it visits): in the
// HierView itself, or in a copy the kernel placed in LDS.
// offz(k) is the offset with level 0 read as 0 (it lives in the caller's array).
struct LevelsOf {
  static constexpr bool kWhole0 = false;
  static constexpr bool kBlockScans = false;
  const HierView& H;
  __device__ __forceinline__ uint32_t off(uint32_t k) const { return H.off[k]; }
  __device__ __forceinline__ uint32_t offz(uint32_t k) const { return k == 0 ? 0u : H.off[k]; }
  __device__ __forceinline__ uint32_t roff(uint32_t k) const { return H.roff[k]; }
  __device__ __forceinline__ uint32_t size(uint32_t k) const { return H.size[k]; }
};
// W0: level 0 (the caller's array) holds whole 16-byte vectors (n % 4 == 0), so its loads need
// no tail select.
// PS: the hierarchy has block scans (HierView::ps), which the lane and group solvers read.
template <bool W0, bool PS = false>
struct LevelsLdsT {
  static constexpr bool kWhole0 = W0;
  static constexpr bool kBlockScans = PS;
  const uint32_t* meta;  // LDS: off (with off[0] = 0), size, roff; kMaxLevels each (lds_meta)
  __device__ __forceinline__ uint32_t off(uint32_t k) const { return meta[k]; }
  __device__ __forceinline__ uint32_t offz(uint32_t k) const { return meta[k]; }
  __device__ __forceinline__ uint32_t size(uint32_t k) const { return meta[kMaxLevels + k]; }
  __device__ __forceinline__ uint32_t roff(uint32_t k) const { return meta[2 * kMaxLevels + k]; }
};
using LevelsLds = LevelsLdsT<false>;
constexpr uint32_t kMetaWords = 3 * kMaxLevels;

// Entry i (< kMetaWords) of the LDS copy that LevelsLds reads.
__device__ __forceinline__ uint32_t lds_meta(const HierView& H, uint32_t i) {
  if (i < kMaxLevels) return i == 0 ? 0u : H.off[i];
  if (i < 2 * kMaxLevels) return H.size[i - kMaxLevels];
  return H.roff[i - 2 * kMaxLevels];
}

// The winner's level-0 index for a group candidate: its key on level 0, else the key (the
// level-0 start of its span) plus the level's offset entry.
template <int RL, class LT>
__device__ __forceinline__ uint32_t cand_index(const dev::Cand& b, const HierView& H, const LT& T) {
  if (b.pay & dev::kDirect) return b.key;
  const uint32_t sk = RL * b.pay;
  const uint32_t j = sk >= 32u ? 0u : (b.key >> sk);
  return b.key + dev::rel_load(H.lr, T.roff(b.pay), j, sk);
}

template <int G, int RL, bool IDX>
__device__ __forceinline__ GroupAnswer solve_cl(uint32_t l, uint32_t e, uint32_t gl,
                                                const HierView& H) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;
  float bv = RMQ_INF;
  dev::Cand b = dev::cand_none();
  const float* cur_v = H.A;
  uint32_t cur_size = H.size[0];
  uint32_t shift = 0, pay = dev::kDirect;
  bool climbing = true;
  for (uint32_t lev = 0; lev + 1 < H.num_levels; ++lev) {
    if (climbing && e - l > 2 * R) {
      const uint32_t nl = (l + R - 1) & ~(R - 1);
      const uint32_t pr = e & ~(R - 1);
      const uint32_t lb = l & ~(R - 1);
#pragma unroll
      for (uint32_t s = gl; s < V; s += G) {
        const uint32_t p = lb + 4 * s;
        if (p + 4 > l && p < nl) {
          const float4 x = dev::load4(cur_v, p, cur_size);
          if constexpr (IDX) dev::cand_fold4(b, x, p, l, nl, shift, pay);
          else bv = dev::fold4(bv, x, p, l, nl);
        }
        const uint32_t p2 = pr + 4 * s;
        if (p2 < e) {
          const float4 x = dev::load4(cur_v, p2, cur_size);
          if constexpr (IDX) dev::cand_fold4(b, x, p2, pr, e, shift, pay);
          else bv = dev::fold4(bv, x, p2, pr, e);
        }
      }
      l = nl >> RL;
      e = pr >> RL;
      cur_v = H.lv + H.off[lev + 1];
      cur_size = H.size[lev + 1];
      shift += RL;
      pay = lev + 1;
    } else {
      climbing = false;
    }
  }
  for (uint32_t p = (l & ~(4 * G - 1)) + 4 * gl; p < e; p += 4 * G) {
    if (p + 4 > l) {
      const float4 x = dev::load4(cur_v, p, cur_size);
      if constexpr (IDX) dev::cand_fold4(b, x, p, l, e, shift, pay);
      else bv = dev::fold4(bv, x, p, l, e);
    }
  }
  if constexpr (IDX) {
    dev::group_cand<G>(b);
    return GroupAnswer{b.v, cand_index<RL>(b, H, LevelsOf{H})};
  } else {
    return GroupAnswer{dev::group_min<G>(bv), dev::kNoIdx};
  }
}

// Fold the part [lo, hi) of one level-k block (or any range inside one block) into b.
template <int G, int RL, bool IDX>
__device__ __forceinline__ void fold_piece(dev::Cand& b, float& bv, const float* vals,
                                           uint32_t size, uint32_t lo, uint32_t hi, uint32_t gl,
                                           uint32_t shift, uint32_t pay) {
  constexpr uint32_t V = (1u << RL) / 4;
  const uint32_t lb = lo & ~((1u << RL) - 1);
#pragma unroll
  for (uint32_t s = gl; s < V; s += G) {
    const uint32_t p = lb + 4 * s;
    if (p + 4 > lo && p < hi) {
      const float4 x = dev::load4(vals, p, size);
      if constexpr (IDX) dev::cand_fold4(b, x, p, lo, hi, shift, pay);
      else bv = dev::fold4(bv, x, p, lo, hi);
    }
  }
}

// Block scans answer a final range only when it is longer than this: a shorter one is one round
// of four vector loads in the lane solver, and at most two rounds of a G-lane group's (G = 8:
// 16 measured 10% ahead of 4G = 32, G = 4 even; 4 lost 4% on short lengths). In the lane
// solver they answer tail pieces only where R >= 16 (a piece of R <= 8 is at most two vectors,
// whose loads also hold the next guard).
constexpr uint32_t kScanFinalMin = 16;

// TOP = false leaves the top level's sparse table out (for kernels short of registers).
template <int G, int RL, bool IDX, bool TOP = true, class LT>
__device__ __forceinline__ GroupAnswer solve_cl_pruned(uint32_t l, uint32_t e, uint32_t gl,
                                                       const HierView& H, const LT& T) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t STEP = 4 * G;
  // K = the highest level whose aligned blocks include one inside [l, e), in O(1) (the
  // reference climbs while the range is longer than 2R; both decompose [l, e) exactly)
  const uint32_t top = H.num_levels - 1;
  uint32_t K = min((31u - static_cast<uint32_t>(__clz(e - l))) / RL, top);
  uint32_t shK = RL * K;
  uint32_t lk = (l >> shK) + ((l & ((1u << shK) - 1u)) != 0u);
  uint32_t ek = e >> shK;
  if (lk >= ek) {
    --K;
    shK -= RL;
    lk = (l >> shK) + ((l & ((1u << shK) - 1u)) != 0u);
    ek = e >> shK;
  }
  const uint32_t offK = T.offz(K);
  const float* vK = K == 0 ? H.A : H.lv + offK;
  const uint32_t szK = T.size(K);
  const uint32_t payK = K == 0 ? dev::kDirect : K;
  const bool left_tail = l < (lk << shK);
  const bool right_tail = (ek << shK) < e;
  // guards are issued together with the final range
  const float gL = left_tail ? vK[lk - 1] : RMQ_INF;
  const float gR = right_tail ? vK[ek] : RMQ_INF;

  // final range: per lane in increasing position, strict `<` keeps the leftmost; a long one on
  // the top level is the sparse table's (one lane looks it up, the group's choice is uniform)
  float bv = RMQ_INF;
  uint32_t bpos = dev::kNoIdx;
  const bool table = TOP && K == top && H.tt != nullptr && ek - lk > dev::kTopScan;
  if (table && gl == 0) dev::top_lookup(H.tt, H.tstride, lk, ek, bv, bpos);
  // block scans: a final range of more than kScanFinalMin entries that is the rest of one
  // R-block and the start of the next is the first's suffix minimum or the second's prefix
  // minimum (one lane reads both, with their level-0 indices)
  const bool scans = LT::kBlockScans && !table && K >= H.ps_lo && ek - lk > kScanFinalMin &&
                     ((ek - 1u) >> RL) == (lk >> RL) + 1u;
  if (scans && gl == 0) {
    const uint2* ps = reinterpret_cast<const uint2*>(H.ps + offK);
    const uint2 x = ps[2u * lk];
    const uint2 y = ps[2u * (ek - 1u) + 1u];
    const bool right = dev::u2f(y.x) < dev::u2f(x.x);
    bv = dev::u2f(right ? y.x : x.x);
    bpos = right ? y.y : x.y;
  }
  const bool direct = table || scans;  // bpos is then a level-0 index
  for (uint32_t p = (lk & ~(STEP - 1)) + 4 * gl; !direct && p < ek; p += STEP) {
    if (p + 4 > lk) {
      const float4 x = dev::load4(vK, p, szK);
      if constexpr (IDX) {
        const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
        for (uint32_t j = 0; j < 4; ++j) {
          // the first element seeds the lane (an all-+inf range answers with its leftmost), but
          // never a NaN: a lane holding one would disagree with its group in the reduction
          const bool t = dev::in_range(p + j, lk, ek) && (xs[j] < bv || (bpos == dev::kNoIdx && xs[j] == xs[j]));
          bv = t ? xs[j] : bv;
          bpos = t ? p + j : bpos;
        }
      } else {
        bv = dev::fold4(bv, x, p, lk, ek);
      }
    }
  }
  dev::Cand b;
  if constexpr (IDX) {
    b.v = bv;
    // a table or block-scan answer is a level-0 index already (a key inside the entry's span,
    // so candidates keep their order); a scanned one is a level-K position
    b.key = bpos == dev::kNoIdx ? dev::kNoIdx : (direct ? bpos : bpos << shK);
    b.pay = bpos == dev::kNoIdx ? dev::kNoIdx : (direct ? dev::kDirect : payK);
    // a direct answer is the first lane's alone: broadcast it instead of reducing
    if (direct) dev::group_first_cand<G>(b);
    else dev::group_cand<G>(b);
    bv = b.v;
  } else {
    bv = direct ? dev::u2f(dev::group_first<G>(dev::f2u(bv))) : dev::group_min<G>(bv);
  }

  // Both tails walk down together (rare for random data): per level the group folds the left
  // piece and the right piece and reduces once. Candidates compare by (value, span start), so
  // the order of the pieces does not matter; bv is group-reduced, so every decision is uniform
  // across the group.
  bool la = left_tail && !(gL > bv);
  bool ra = right_tail && gR < bv;
  for (int k = static_cast<int>(K) - 1; k >= 0 && (la || ra); --k) {
    const uint32_t uk = static_cast<uint32_t>(k);
    const uint32_t sk = RL * uk;
    const uint32_t ok = T.offz(uk);
    const float* vk = k == 0 ? H.A : H.lv + ok;
    const uint32_t szk = T.size(uk);
    const uint32_t pay = k == 0 ? dev::kDirect : uk;
    const uint32_t lo = (l >> sk) + ((l & ((1u << sk) - 1u)) != 0u);
    const uint32_t hi = e >> sk;
    // the guards of the level below, loaded with this level's pieces rather than after them
    const bool nl = la && k != 0 && l < (lo << sk);
    const bool nr = ra && k != 0 && (hi << sk) < e;
    const float g_next_l = nl ? vk[lo - 1] : RMQ_INF;
    const float g_next_r = nr ? vk[hi] : RMQ_INF;
    const bool by_scans = LT::kBlockScans && uk >= H.ps_lo;
    if (by_scans) {
      // a piece is the rest of its block (left) or the block's start (right): one block-scan
      // entry each, read by one lane (an empty piece, on a block boundary, reads nothing)
      if (gl == 0) {
        const uint2* ps = reinterpret_cast<const uint2*>(H.ps + ok);
        if (la && (lo & (R - 1u)) != 0u) {
          const uint2 x = ps[2u * lo];
          if constexpr (IDX) dev::cand_take(b, dev::u2f(x.x), x.y, dev::kDirect, true);
          else bv = dev::fmin_sel(bv, dev::u2f(x.x));
        }
        if (ra && (hi & (R - 1u)) != 0u) {
          const uint2 y = ps[2u * (hi - 1u) + 1u];
          if constexpr (IDX) dev::cand_take(b, dev::u2f(y.x), y.y, dev::kDirect, true);
          else bv = dev::fmin_sel(bv, dev::u2f(y.x));
        }
      }
    } else {
      if (la) fold_piece<G, RL, IDX>(b, bv, vk, szk, lo, (lo + R - 1) & ~(R - 1), gl, sk, pay);
      if (ra) fold_piece<G, RL, IDX>(b, bv, vk, szk, hi & ~(R - 1), hi, gl, sk, pay);
    }
    // the group held one best before this level; by block scans only the first lane moved
    if constexpr (IDX) {
      if (by_scans) dev::group_first_cand<G>(b);
      else dev::group_cand<G>(b);
      bv = b.v;
    } else {
      bv = by_scans ? dev::u2f(dev::group_first<G>(dev::f2u(bv))) : dev::group_min<G>(bv);
    }
    la = nl && !(g_next_l > bv);  // deeper left pieces can still win
    ra = nr && g_next_r < bv;
  }
  if constexpr (IDX) return GroupAnswer{b.v, cand_index<RL>(b, H, T)};
  else return GroupAnswer{bv, dev::kNoIdx};
}

// P selects the pruned evaluation order (default) or the full one.
template <int G, int RL, bool P, bool IDX, bool TOP = true, class LT>
__device__ __forceinline__ GroupAnswer solve_group(uint32_t l, uint32_t e, uint32_t gl,
                                                   const HierView& H, const LT& T) {
  if constexpr (P) return solve_cl_pruned<G, RL, IDX, TOP>(l, e, gl, H, T);
  else return solve_cl<G, RL, IDX>(l, e, gl, H);
}

// Value-only answers must carry the bit pattern of A[leftmost argmin]. A plain minimum can
// only differ from that when the minimum is zero (+0 and -0 compare equal): the zero's sign is
// taken from the array's zero signs, and only when the array holds both is the query answered
// again with index tracking.
template <int G, int RL, bool P, bool TOP = true, class LT>
__device__ __forceinline__ float value_exact(uint32_t l, uint32_t e, uint32_t gl,
                                             const HierView& H, const LT& T, uint32_t zeros) {
  GroupAnswer a = solve_group<G, RL, P, false, TOP>(l, e, gl, H, T);
  if (a.v == 0.0f && !dev::zero_known(zeros, a.v)) a = solve_group<G, RL, P, true, TOP>(l, e, gl, H, T);
  return a.v;
}

// Alg 16 (xxx_query, src/xxx.h:124-200): every lane owns one query; a G-lane group serves
// its members' queries one after another (the reference's ballot work queue, here a fixed
// round over the group's lanes with a ds_bpermute broadcast). Queries are read and answers
// written with one coalesced access per wave.
template <int G, int RL, bool P, bool IDX, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_xxx(const uint2* __restrict__ Q, uint32_t q,
                                                      float* __restrict__ out_v,
                                                      uint32_t* __restrict__ out_i, HierView H) {
  __shared__ uint32_t s_meta[kMetaWords];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LevelsLdsT<false, PS> T{s_meta};
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gl = lane & (G - 1);
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool valid = qi < q;
  const uint32_t zeros = *H.zeros;  // zero signs and the NaN bit of level 0
  uint32_t l = 0, e = 0;
  if (valid) {
    const uint2 p = dev::load_query(Q, qi);
    l = p.x;
    e = p.y + 1;
  }
  float my_v = RMQ_INF;
  uint32_t my_i = dev::kNoIdx;
  const int gbase = static_cast<int>(lane & ~(G - 1));
  for (int k = 0; k < G; ++k) {
    const int src = gbase + k;
    const uint32_t lk = static_cast<uint32_t>(__shfl(static_cast<int>(l), src));
    const uint32_t ek = static_cast<uint32_t>(__shfl(static_cast<int>(e), src));
    const int vk = __shfl(static_cast<int>(valid), src);
    if (vk) {
      float v;
      uint32_t i = dev::kNoIdx;
      if constexpr (IDX) {
        const GroupAnswer a = solve_group<G, RL, P, true>(lk, ek, gl, H, T);
        v = a.v;
        i = a.idx;
      } else {
        v = value_exact<G, RL, P>(lk, ek, gl, H, T, zeros);
      }
      if (gl == static_cast<uint32_t>(k)) {
        my_v = v;
        my_i = i;
      }
    }
  }
  if (valid) {
    if (dev::nan_at_l(zeros, H.A, l, my_v)) my_i = l;
    if (out_v) out_v[qi] = my_v;
    if (IDX && out_i) out_i[qi] = my_i;
  }
}

// Alg 24 (xxx_query_multi_load, src/xxx.h:470-528): one G-lane group per query, every lane of
// the group reads the same query; grid-stride over queries in units of groups (the
// reference strides by the thread count, Appendix-A quirk not kept).
template <int G, int RL, bool P, bool IDX, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_multi(const uint2* __restrict__ Q, uint32_t q,
                                                        float* __restrict__ out_v,
                                                        uint32_t* __restrict__ out_i,
                                                        HierView H) {
  __shared__ uint32_t s_meta[kMetaWords];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LevelsLdsT<false, PS> T{s_meta};
  const uint32_t gl = threadIdx.x & (G - 1);
  const uint64_t grp = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / G;
  const uint64_t ngrp = static_cast<uint64_t>(gridDim.x) * kBlock / G;
  const uint32_t zeros = *H.zeros;
  for (uint64_t qi = grp; qi < q; qi += ngrp) {
    const uint2 p = dev::load_query(Q, qi);
    float v;
    uint32_t i = dev::kNoIdx;
    if constexpr (IDX) {
      const GroupAnswer a = solve_group<G, RL, P, true>(p.x, p.y + 1, gl, H, T);
      v = a.v;
      i = a.idx;
    } else {
      v = value_exact<G, RL, P>(p.x, p.y + 1, gl, H, T, zeros);
    }
    if (gl == 0) {
      if (dev::nan_at_l(zeros, H.A, p.x, v)) i = p.x;
      if (out_v) out_v[qi] = v;
      if (IDX && out_i) out_i[qi] = i;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Alg 20 (kernel_hierarchical_vector_load, src/xxx.h:883-930): one lane per query, 16-byte
// loads, R = 4*2^c. Climb while r - l > R and the aligned bounds do not cross; fold
// [l, ceil_R(l)) and [floor_R(r), r); finish with a 16-byte scan of [l, r). (The reference
// also re-reads the whole block when l is already aligned; that block is covered by the
// next level, so it is skipped here — same answer, fewer loads.)
template <int RL, bool IDX>
__device__ __forceinline__ GroupAnswer solve_vl(uint32_t l, uint32_t e, const HierView& H) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;
  float bv = RMQ_INF;
  dev::Cand b = dev::cand_none();
  const float* cur_v = H.A;
  uint32_t cur_size = H.size[0];
  uint32_t shift = 0, pay = dev::kDirect;
  bool climbing = true;
  for (uint32_t lev = 0; lev + 1 < H.num_levels; ++lev) {
    if (climbing && e - l > R) {
      const uint32_t nl = (l + R - 1) & ~(R - 1);
      const uint32_t pr = e & ~(R - 1);
      if (pr >= nl) {
        const uint32_t lb = l & ~(R - 1);
        float4 xl[V], xr[V];
#pragma unroll
        for (uint32_t s = 0; s < V; ++s) {  // both blocks' loads in flight together
          const uint32_t p = lb + 4 * s;
          xl[s] = (p + 4 > l && p < nl) ? dev::load4(cur_v, p, cur_size)
                                        : make_float4(RMQ_INF, RMQ_INF, RMQ_INF, RMQ_INF);
          const uint32_t p2 = pr + 4 * s;
          xr[s] = p2 < e ? dev::load4(cur_v, p2, cur_size)
                         : make_float4(RMQ_INF, RMQ_INF, RMQ_INF, RMQ_INF);
        }
#pragma unroll
        for (uint32_t s = 0; s < V; ++s) {
          if constexpr (IDX) {
            dev::cand_fold4(b, xl[s], lb + 4 * s, l, nl, shift, pay);
            dev::cand_fold4(b, xr[s], pr + 4 * s, pr, e, shift, pay);
          } else {
            bv = dev::fold4(bv, xl[s], lb + 4 * s, l, nl);
            bv = dev::fold4(bv, xr[s], pr + 4 * s, pr, e);
          }
        }
        l = nl >> RL;
        e = pr >> RL;
        cur_v = H.lv + H.off[lev + 1];
        cur_size = H.size[lev + 1];
        shift += RL;
        pay = lev + 1;
      } else {
        climbing = false;
      }
    } else {
      climbing = false;
    }
  }
  for (uint32_t p = l & ~3u; p < e; p += 4) {
    const float4 x = dev::load4(cur_v, p, cur_size);
    if constexpr (IDX) dev::cand_fold4(b, x, p, l, e, shift, pay);
    else bv = dev::fold4(bv, x, p, l, e);
  }
  if constexpr (IDX) return GroupAnswer{b.v, cand_index<RL>(b, H, LevelsOf{H})};
  else return GroupAnswer{bv, dev::kNoIdx};
}

// A 16-byte vector of level k at p (p % 4 == 0, p < size), with no branch: upper levels sit in
// the level buffer, padded to 32 entries past each level's end, so their vectors are whole; on
// level 0 (the caller's array of exactly n floats) the one vector that would run past n, at
// p = n & ~3, is read from the hierarchy's copy of it, padded with +inf. Where the level table
// type says that n % 4 == 0 (LT::kWhole0), there is no such vector and no select.
template <class LT>
__device__ __forceinline__ float4 load_level(const float* __restrict__ v, uint32_t p, uint32_t size, uint32_t k,
                                             const float* __restrict__ tail) {
  if constexpr (LT::kWhole0) return *reinterpret_cast<const float4*>(v + p);
  const float* src = (k != 0 || p + 4u <= size) ? v + p : tail;
  return *reinterpret_cast<const float4*>(src);
}

// Lane-per-query pruned evaluation (the pieces of alg 20, read in the order of
// solve_cl_pruned). The final range is taken at the highest level K whose
// aligned blocks include one inside [l, e), found in O(1) from the length (any level whose
// range is nonempty yields the same answer; this one leaves at most 2R - 2 entries to scan
// below the top level). Sides need no position keys: every left piece lies left of all that was
// read before it, and every right piece right of it, so a left element wins ties (folded right
// to left with <=) and a right element must be strictly smaller (folded left to right with <).
// In the final range, folded left to right with <, the first entry seeds the answer, so a range
// of +inf answers with its leftmost entry.
//
// The solver in three resumable parts, so that a kernel can hold a query between levels:
// vl_start reads the final range and its guards, vl_step walks both tails down one level,
// vl_finish resolves the winner's index. A query needs vl_step while la or ra holds.
// The lane solver's winner payload for position pos of level k (blocks of 2^sk level-0
// elements, sk = R_log k): kDirect | pos on level 0; else the level-0 span start pos << sk with
// bit sk - 1 set, so that the level is read back from the lowest set bit. pay(pos) =
// base + (pos << sk) with base = 2^((sk - 1) mod 32), which is kDirect on level 0. The solver's
// levels have sk <= 30 (its final range sits at most floor(30 / R_log) levels up).
__device__ __forceinline__ uint32_t pay_base(uint32_t sk) { return 1u << ((sk + 31u) & 31u); }

template <int RL, class LT>
__device__ __forceinline__ uint32_t pay_index(uint32_t pay, const HierView& H, const LT& T) {
  if (pay & dev::kDirect) return pay & ~dev::kDirect;
  const uint32_t sk = static_cast<uint32_t>(__builtin_ctz(pay)) + 1u;
  const uint32_t s0 = pay & ~((1u << sk) - 1u);
  return s0 + dev::rel_load(H.lr, T.roff(sk / RL), s0 >> sk, sk);
}


struct VlState {
  uint32_t l, e;
  float bv;
  uint32_t bpay;  // winner (pay_base, pay_index)
  int k;          // the level the next step reads (K - 1 down to 0)
  bool la, ra;    // a tail's next piece can still win
};

// TOP: long final ranges on the top level come from its sparse table when the hierarchy has
// one; EARLY (R <= 8): a short final range's vectors are all loaded before any is folded.
// The level of [l, e)'s final range and its bounds there (lk = ceil(l / R^K), ek = e >> shK).
struct VlLevel {
  uint32_t K, shK, lk, ek;
};

template <int RL>
__device__ __forceinline__ VlLevel vl_level(uint32_t l, uint32_t e, uint32_t top) {
  uint32_t K = min((31u - static_cast<uint32_t>(__clz(e - l))) / RL, top);
  uint32_t shK = RL * K;
  uint32_t lk = (l >> shK) + ((l & ((1u << shK) - 1u)) != 0u);
  uint32_t ek = e >> shK;
  if (lk >= ek) {  // K >= 1 here, and the level below holds a block inside [l, e)
    --K;
    shK -= RL;
    lk = (l >> shK) + ((l & ((1u << shK) - 1u)) != 0u);
    ek = e >> shK;
  }
  return VlLevel{K, shK, lk, ek};
}

template <int RL, bool IDX, bool TOP, class LT, bool EARLY = TOP>
__device__ __forceinline__ VlState vl_start_at(uint32_t l, uint32_t e, const VlLevel& lv, const HierView& H,
                                               const LT& T) {
  const uint32_t top = H.num_levels - 1;
  const uint32_t K = lv.K, shK = lv.shK, lk = lv.lk, ek = lv.ek;
  const float* vK = K == 0 ? H.A : H.lv + T.offz(K);
  const uint32_t szK = T.size(K);
  const bool left_tail = l < (lk << shK);
  const bool right_tail = (ek << shK) < e;
  float bv = RMQ_INF;
  const uint32_t pb = pay_base(shK);
  uint32_t bpay = pb + (lk << shK);
  float gL = RMQ_INF, gR = RMQ_INF;
  if (TOP && K == top && H.tt != nullptr && ek - lk > dev::kTopScan) {
    // a long final range on the top level: the sparse table's leftmost minimum
    uint32_t idx0;  // the level-0 index of the minimum: a direct payload
    dev::top_lookup(H.tt, H.tstride, lk, ek, bv, idx0);
    bpay = dev::kDirect | idx0;
    if (left_tail) gL = vK[lk - 1u];
    if (right_tail) gR = vK[ek];
  } else if (LT::kBlockScans && RL >= 4 && K >= H.ps_lo && ek - lk > kScanFinalMin &&
             ((ek - 1u) >> RL) == (lk >> RL) + 1u) {
    // a final range of more than one round of four vectors that is the rest of one R-block and
    // the start of the next: the suffix minimum from lk and the prefix minimum up to ek - 1,
    // with their level-0 indices
    const uint2* ps = reinterpret_cast<const uint2*>(H.ps + T.offz(K));
    const uint2 x = ps[2u * lk];
    const uint2 y = ps[2u * (ek - 1u) + 1u];
    if (left_tail) gL = vK[lk - 1u];
    if (right_tail) gR = vK[ek];
    const bool right = dev::u2f(y.x) < dev::u2f(x.x);
    bv = dev::u2f(right ? y.x : x.x);
    bpay = dev::kDirect | (right ? y.y : x.y);
  } else {
    // guards lk - 1 and ek: in the first / last vector the scan reads unless 4-aligned
    uint32_t p = lk & ~3u;
    // R <= 8: a range below the top (at most 2R - 2 entries) or on a top level with a table (at
    // most kTopScan) spans at most NV vectors, all loaded before any is folded; a wavefront pays
    // for its longest range either way, so this trades a dependent loop for parallel loads.
    // (Not in kernels without the table path: the batched kernel has no registers to spare.)
    constexpr uint32_t kScanMax = 2 * (1u << RL) - 2 > dev::kTopScan ? 2 * (1u << RL) - 2 : dev::kTopScan;
    constexpr uint32_t NV = (kScanMax + 3 + 3) / 4;
    if (EARLY && RL <= 3 && ek - lk <= kScanMax) {
      const float ga = left_tail && (lk & 3u) == 0 ? vK[lk - 1u] : RMQ_INF;
      const float gb = right_tail && (ek & 3u) == 0 ? vK[ek] : RMQ_INF;
      float4 xv[NV];
#pragma unroll
      for (uint32_t v = 0; v < NV; ++v)
        if (p + 4u * v < ek) xv[v] = load_level<LT>(vK, p + 4u * v, szK, K, H.tail);
      if (left_tail) gL = (lk & 3u) ? dev::pick4(xv[0], (lk - 1u) & 3u) : ga;
      // index mode: the winner's offset from p (a constant per element), its payload once after
      uint32_t bo = lk - p;
#pragma unroll
      for (uint32_t v = 0; v < NV; ++v) {
        const uint32_t pv = p + 4u * v;
        if (pv < ek) {
          if constexpr (IDX) {
            const float xs[4] = {xv[v].x, xv[v].y, xv[v].z, xv[v].w};
#pragma unroll
            for (uint32_t j = 0; j < 4; ++j) {
              const bool t = dev::in_range(pv + j, lk, ek) && xs[j] < bv;
              bv = t ? xs[j] : bv;
              bo = t ? 4u * v + j : bo;
            }
          } else {
            bv = dev::fold4(bv, xv[v], pv, lk, ek);
          }
          if (right_tail && (ek & 3u) != 0 && ek - pv < 4u) gR = dev::pick4(xv[v], ek & 3u);
        }
      }
      if constexpr (IDX) bpay = pb + ((p + bo) << shK);
      if (right_tail && (ek & 3u) == 0) gR = gb;
      return VlState{l, e, bv, bpay, static_cast<int>(K) - 1, left_tail && gL <= bv, right_tail && gR < bv};
    }
    if constexpr (RL >= 4) {
      // R >= 16: a final range below the top spans up to 2R - 2 entries (16 vectors at R = 32);
      // four vectors are loaded before any of them is folded, a quarter of the round trips
      float4 last = make_float4(RMQ_INF, RMQ_INF, RMQ_INF, RMQ_INF);
      for (bool first = true;; first = false) {
        float4 xv[4];
#pragma unroll
        for (uint32_t v = 0; v < 4; ++v)
          if (p + 4u * v < ek) xv[v] = load_level<LT>(vK, p + 4u * v, szK, K, H.tail);
        if (first && left_tail) gL = (lk & 3u) ? dev::pick4(xv[0], (lk - 1u) & 3u) : vK[lk - 1u];
#pragma unroll
        for (uint32_t v = 0; v < 4; ++v) {
          const uint32_t pv = p + 4u * v;
          if (pv < ek) {
            const float4 x = xv[v];
            if constexpr (IDX) {
              const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
              for (uint32_t j = 0; j < 4; ++j) {
                const bool t = dev::in_range(pv + j, lk, ek) && xs[j] < bv;
                bv = t ? xs[j] : bv;
                bpay = t ? pb + ((pv + j) << shK) : bpay;
              }
            } else {
              bv = dev::fold4(bv, x, pv, lk, ek);
            }
            last = x;
          }
        }
        p += 16;
        if (p >= ek) break;
      }
      if (right_tail) gR = (ek & 3u) ? dev::pick4(last, ek & 3u) : vK[ek];
    } else {
      float4 x = load_level<LT>(vK, p, szK, K, H.tail);
      if (left_tail) gL = (lk & 3u) ? dev::pick4(x, (lk - 1u) & 3u) : vK[lk - 1u];
      for (;;) {
        if constexpr (IDX) {
          const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
          for (uint32_t j = 0; j < 4; ++j) {
            const bool t = dev::in_range(p + j, lk, ek) && xs[j] < bv;
            bv = t ? xs[j] : bv;
            bpay = t ? pb + ((p + j) << shK) : bpay;
          }
        } else {
          bv = dev::fold4(bv, x, p, lk, ek);
        }
        p += 4;
        if (p >= ek) break;
        x = load_level<LT>(vK, p, szK, K, H.tail);
      }
      if (right_tail) gR = (ek & 3u) ? dev::pick4(x, ek & 3u) : vK[ek];
    }
  }
  return VlState{l, e, bv, bpay, static_cast<int>(K) - 1, left_tail && gL <= bv, right_tail && gR < bv};
}

template <int RL, bool IDX, bool TOP, class LT, bool EARLY = TOP>
__device__ __forceinline__ VlState vl_start(uint32_t l, uint32_t e, const HierView& H, const LT& T) {
  return vl_start_at<RL, IDX, TOP, LT, EARLY>(l, e, vl_level<RL>(l, e, H.num_levels - 1), H, T);
}

// One level of both tails. Any interleaving of the two sides is exact: a left element wins ties
// whenever it is folded, a right one never does, and each guard test compares with the best so
// far; both sides' loads of a level are in flight together.
template <int RL, bool IDX, class LT>
__device__ __forceinline__ void vl_step(VlState& st, const HierView& H, const LT& T) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;
  const uint32_t l = st.l, e = st.e;
  float bv = st.bv;
  uint32_t bpay = st.bpay;
  bool la = st.la, ra = st.ra;
  const int k = st.k;
  const uint32_t uk = static_cast<uint32_t>(k);
  const uint32_t sk = RL * uk;
  const uint32_t ok = T.offz(uk);
  const float* vk = k == 0 ? H.A : H.lv + ok;
  const uint32_t szk = T.size(uk);
  const uint32_t pb = pay_base(sk);
  const uint32_t lo = (l >> sk) + ((l & ((1u << sk) - 1u)) != 0u);  // left piece [lo, roundup_R(lo))
  const uint32_t lbase = lo & ~(R - 1u);
  const uint32_t hi = e >> sk;                                       // right piece [rlo, hi)
  const uint32_t rlo = hi & ~(R - 1u);
  // R <= 8: both pieces' vectors are loaded before either is folded; wider blocks load per
  // side (holding 2 x R/4 vectors would cost the occupancy the overlap is meant to buy)
  constexpr bool kEarly = V <= 2;
  constexpr uint32_t VE = kEarly ? V : 1;
  // each vector is read below only under the condition it was loaded under
  float4 xl[VE], xr[VE];
  // the next guards, where they lie outside the vectors read (4-aligned piece ends), are loaded
  // with them rather than after the fold: same entries, same conditions, one round trip less
  const bool lnext = la && k != 0 && l < (lo << sk);
  const bool rnext = ra && k != 0 && (hi << sk) < e;
  if constexpr (LT::kBlockScans && RL >= 4) {
    if (uk >= H.ps_lo) {
      // each piece is one block-scan entry: the left piece [lo, block end) a suffix minimum, the
      // right piece [block start, hi) a prefix minimum; the next guards are loaded with them
      const uint2* ps = reinterpret_cast<const uint2*>(H.ps + ok);
      const bool lp = la && (lo & (R - 1u)) != 0u;
      const bool rp = ra && (hi & (R - 1u)) != 0u;
      const uint2 none = make_uint2(dev::f2u(RMQ_INF), 0u);
      const uint2 x = lp ? ps[2u * lo] : none;
      const uint2 y = rp ? ps[2u * (hi - 1u) + 1u] : none;
      const float gl = lnext ? vk[lo - 1u] : RMQ_INF;
      const float gr = rnext ? vk[hi] : RMQ_INF;
      if (lp && dev::u2f(x.x) <= bv) {
        bv = dev::u2f(x.x);
        bpay = dev::kDirect | x.y;
      }
      if (rp && dev::u2f(y.x) < bv) {
        bv = dev::u2f(y.x);
        bpay = dev::kDirect | y.y;
      }
      st.bv = bv;
      st.bpay = bpay;
      st.la = lnext && !(gl > bv);
      st.ra = rnext && gr < bv;
      st.k = k - 1;
      return;
    }
  }
  const float gl0 = lnext && (lo & 3u) == 0 ? vk[lo - 1u] : RMQ_INF;
  const float gr0 = rnext && (hi & 3u) == 0 ? vk[hi] : RMQ_INF;
  if constexpr (kEarly) {
#pragma unroll
    for (uint32_t s = 0; s < V; ++s) {
      const uint32_t pl = lbase + 4u * s, pr = rlo + 4u * s;
      if (la && (lo & (R - 1u)) != 0u && pl + 4u > lo) xl[s] = load_level<LT>(vk, pl, szk, uk, H.tail);
      if (ra && pr < hi) xr[s] = load_level<LT>(vk, pr, szk, uk, H.tail);
    }
  }
  // index mode: the winning element's slot (a constant per element; R + slot on the right),
  // its payload formed once at the end of the step
  constexpr uint32_t kNoSlot = 2u * R;
  uint32_t wslot = kNoSlot;
  // folds one vector of the left piece (slot s: right to left, `<=`) / of the right piece
  // (left to right, strict `<`), noting the next guard where the vector holds it
  auto fold_left = [&](const float4 x, uint32_t s, float& g) {
    const uint32_t p = lbase + 4u * s;
    if (p <= lo - 1u) g = dev::pick4(x, (lo - 1u) & 3u);  // next guard, lo - 1
    const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (int j = 3; j >= 0; --j) {
      const bool t = p + static_cast<uint32_t>(j) >= lo && xs[j] <= bv;
      bv = t ? xs[j] : bv;
      if constexpr (IDX) wslot = t ? 4u * s + static_cast<uint32_t>(j) : wslot;
    }
  };
  auto fold_right = [&](const float4 x, uint32_t s, float& g) {
    const uint32_t p = rlo + 4u * s;
    if (hi - p < 4u) g = dev::pick4(x, hi & 3u);  // next guard, hi
    const float xs[4] = {x.x, x.y, x.z, x.w};
#pragma unroll
    for (uint32_t j = 0; j < 4; ++j) {
      const bool t = p + j < hi && xs[j] < bv;
      bv = t ? xs[j] : bv;
      if constexpr (IDX) wslot = t ? R + 4u * s + j : wslot;
    }
  };
  // R >= 16 (V >= 4): a piece's vectors are loaded four at a time before they are folded
  constexpr bool kGroups = !kEarly && V % 4u == 0;
  if (la) {
    float g = RMQ_INF;
    if constexpr (kGroups) {
#pragma unroll
      for (int s0 = static_cast<int>(V) - 4; s0 >= 0; s0 -= 4) {
        const uint32_t pg = lbase + 4u * static_cast<uint32_t>(s0);
        if ((lo & (R - 1u)) != 0u && pg + 16u > lo) {
          float4 xg[4];
#pragma unroll
          for (uint32_t v = 0; v < 4; ++v)
            if (pg + 4u * v + 4u > lo) xg[v] = load_level<LT>(vk, pg + 4u * v, szk, uk, H.tail);
#pragma unroll
          for (int v = 3; v >= 0; --v)
            if (pg + 4u * static_cast<uint32_t>(v) + 4u > lo) fold_left(xg[v], static_cast<uint32_t>(s0 + v), g);
        }
      }
    } else {
#pragma unroll
      for (int s = static_cast<int>(V) - 1; s >= 0; --s) {
        const uint32_t p = lbase + 4u * static_cast<uint32_t>(s);
        if ((lo & (R - 1u)) != 0u && p + 4u > lo)
          fold_left(kEarly ? xl[kEarly ? s : 0] : load_level<LT>(vk, p, szk, uk, H.tail), static_cast<uint32_t>(s), g);
      }
    }
    if (!lnext) {
      la = false;
    } else {
      if ((lo & 3u) == 0) g = gl0;
      la = !(g > bv);
    }
  }
  if (ra) {
    float g = RMQ_INF;
    if constexpr (kGroups) {
#pragma unroll
      for (uint32_t s0 = 0; s0 < V; s0 += 4) {
        const uint32_t pg = rlo + 4u * s0;
        if (pg < hi) {
          float4 xg[4];
#pragma unroll
          for (uint32_t v = 0; v < 4; ++v)
            if (pg + 4u * v < hi) xg[v] = load_level<LT>(vk, pg + 4u * v, szk, uk, H.tail);
#pragma unroll
          for (uint32_t v = 0; v < 4; ++v)
            if (pg + 4u * v < hi) fold_right(xg[v], s0 + v, g);
        }
      }
    } else {
#pragma unroll
      for (uint32_t s = 0; s < V; ++s) {
        const uint32_t p = rlo + 4u * s;
        if (p < hi) fold_right(kEarly ? xr[kEarly ? s : 0] : load_level<LT>(vk, p, szk, uk, H.tail), s, g);
      }
    }
    if (!rnext) {
      ra = false;
    } else {
      if ((hi & 3u) == 0) g = gr0;
      ra = g < bv;
    }
  }
  if constexpr (IDX) {
    if (wslot != kNoSlot) bpay = pb + ((wslot < R ? lbase + wslot : rlo + (wslot - R)) << sk);
  }
  st.bv = bv;
  st.bpay = bpay;
  st.la = la;
  st.ra = ra;
  st.k = k - 1;
}

template <int RL, bool IDX, class LT>
__device__ __forceinline__ GroupAnswer vl_finish(const VlState& st, const HierView& H, const LT& T) {
  if constexpr (IDX) return GroupAnswer{st.bv, pay_index<RL>(st.bpay, H, T)};
  else return GroupAnswer{st.bv, dev::kNoIdx};
}

template <int RL, bool IDX, bool TOP = true, class LT, bool EARLY = TOP>
__device__ __forceinline__ GroupAnswer solve_vl_pruned(uint32_t l, uint32_t e, const HierView& H, const LT& T) {
  VlState st = vl_start<RL, IDX, TOP, LT, EARLY>(l, e, H, T);
  while (st.la || st.ra) vl_step<RL, IDX>(st, H, T);
  return vl_finish<RL, IDX>(st, H, T);
}

template <int RL, bool P, bool IDX, class LT>
__device__ __forceinline__ GroupAnswer solve_lane(uint32_t l, uint32_t e, const HierView& H, const LT& T) {
  if constexpr (P) return solve_vl_pruned<RL, IDX>(l, e, H, T);
  else return solve_vl<RL, IDX>(l, e, H);
}

// CPU BASE's NaN rule for a lane's query (dev::nan_at_l): a NaN at A[l] is stored as the answer
// (value A[l], index l) and the query needs no solver. A uniform branch on NaN-free arrays.
template <bool IDX>
__device__ __forceinline__ bool emit_nan_at_l(uint32_t flags, const float* __restrict__ A, uint32_t l, uint64_t qi,
                                              float* __restrict__ out_v, uint32_t* __restrict__ out_i) {
  float v;
  if (!dev::nan_at_l(flags, A, l, v)) return false;
  if (out_v) out_v[qi] = v;
  if (IDX && out_i) out_i[qi] = l;
  return true;
}

template <int RL, bool P, bool IDX, bool W0 = false, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_vl(const uint2* __restrict__ Q, uint32_t q,
                                                     float* __restrict__ out_v,
                                                     uint32_t* __restrict__ out_i, HierView H) {
  // level offsets and sizes in LDS: the solver looks one up per level it visits
  __shared__ uint32_t s_meta[kMetaWords];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LevelsLdsT<W0, PS> T{s_meta};
  const uint32_t zeros = *H.zeros;
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= q) return;
  const uint2 p = dev::load_query(Q, qi);
  if (emit_nan_at_l<IDX>(zeros, H.A, p.x, qi, out_v, out_i)) return;
  GroupAnswer a = solve_lane<RL, P, IDX>(p.x, p.y + 1, H, T);
  if constexpr (!IDX) {
    if (a.v == 0.0f && !dev::zero_known(zeros, a.v)) a = solve_lane<RL, P, true>(p.x, p.y + 1, H, T);
  }
  if (out_v) out_v[qi] = a.v;
  if (IDX && out_i) out_i[qi] = a.idx;
}

// Answer of a finished lane-solver state, written to query qi (value mode re-solves a zero of
// unknown sign).
template <int RL, bool IDX, class LT>
__device__ __forceinline__ void vl_emit(const VlState& st, uint64_t qi, const HierView& H, const LT& T,
                                        uint32_t zeros, float* __restrict__ out_v,
                                        uint32_t* __restrict__ out_i) {
  GroupAnswer a = vl_finish<RL, IDX>(st, H, T);
  if constexpr (!IDX) {
    if (a.v == 0.0f && !dev::zero_known(zeros, a.v)) a = solve_vl_pruned<RL, true>(st.l, st.e, H, T);
  }
  if (out_v) out_v[qi] = a.v;
  if (IDX && out_i) out_i[qi] = a.idx;
}

// The pruned lane solver with its descents packed. k_query_vl's wavefront walks the tails until
// its deepest lane is done, though most queries need at most one level (lognormal lengths at
// R = 4: 0.7 levels per query on average, 3.8 for the deepest of 64). Here every lane reads its
// final range and walks one level; the queries still descending are written to LDS, and after
// the workgroup's barrier the first threads take them up in wavefront order, so the deep walks
// run on as few wavefronts as there are such queries in groups of 64. A lane answers its query
// as soon as it is done. Where few queries descend (uniform lengths) or most descend only a
// little (short lengths), the parking, the barrier and the parked queries' scattered stores
// are not repaid: a few percent slower there.
template <int RL, bool IDX, bool W0 = false, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_vlc(const uint2* __restrict__ Q, uint32_t q,
                                                      float* __restrict__ out_v,
                                                      uint32_t* __restrict__ out_i, HierView H) {
  constexpr uint32_t kWaves = kBlock / 64;
  __shared__ uint32_t s_meta[kMetaWords];
  __shared__ uint4 s_st[kBlock];   // l, e, best value bits, best payload
  __shared__ uint2 s_aux[kBlock];  // next level | la << 8 | ra << 9, query index
  __shared__ uint32_t s_cnt[kWaves];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LevelsLdsT<W0, PS> T{s_meta};
  const uint32_t zeros = *H.zeros;
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  bool need = false;
  VlState st{};
  if (qi < q) {
    const uint2 p = dev::load_query(Q, qi);
    if (!emit_nan_at_l<IDX>(zeros, H.A, p.x, qi, out_v, out_i)) {
      st = vl_start<RL, IDX, true>(p.x, p.y + 1, H, T);
      if (st.la || st.ra) vl_step<RL, IDX>(st, H, T);
      need = st.la || st.ra;  // k >= 0 whenever a tail is open
      if (!need) vl_emit<RL, IDX>(st, qi, H, T, zeros, out_v, out_i);
    }
  }
  const uint64_t m = __ballot(need);
  if (need) {
    const uint32_t slot = w * 64u + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(m >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(m), 0u));
    s_st[slot] = make_uint4(st.l, st.e, dev::f2u(st.bv), st.bpay);
    s_aux[slot] = make_uint2(static_cast<uint32_t>(st.k) | (st.la ? 0x100u : 0u) | (st.ra ? 0x200u : 0u),
                             static_cast<uint32_t>(qi));
  }
  if ((threadIdx.x & 63u) == 0u) s_cnt[w] = static_cast<uint32_t>(__popcll(m));
  __syncthreads();
  uint32_t c[kWaves], total = 0;
#pragma unroll
  for (uint32_t v = 0; v < kWaves; ++v) {
    c[v] = s_cnt[v];
    total += c[v];
  }
  // thread j takes the j-th parked query: wave v's entries follow those of waves 0..v-1 (one
  // pass: at most kBlock are parked)
  for (uint32_t j = threadIdx.x; j < total; j += kBlock) {
    uint32_t r = j, slot = kBlock;
#pragma unroll
    for (uint32_t v = 0; v < kWaves; ++v) {
      if (slot == kBlock) {
        if (r < c[v]) slot = v * 64u + r;
        else r -= c[v];
      }
    }
    const uint4 a = s_st[slot];
    const uint2 b = s_aux[slot];
    VlState ps{a.x, a.y, dev::u2f(a.z), a.w, static_cast<int>(b.x & 0xFFu), (b.x & 0x100u) != 0u,
               (b.x & 0x200u) != 0u};
    while (ps.la || ps.ra) vl_step<RL, IDX>(ps, H, T);
    vl_emit<RL, IDX>(ps, b.y, H, T, zeros, out_v, out_i);
  }
}

// ---------------------------------------------------------------------------------------
// Two queries per lane (RMQ_BUILD_PAIRED). A query answered from the top-level table is a
// chain of two dependent reads (its pair, then the two table entries and its guards); a
// wavefront at full occupancy waits on them. Here each lane takes two adjacent queries (one
// 16-byte pair load) and issues both queries' table entries and guards before using either, so
// twice the reads are in flight per wavefront. A query that is not answered from the table
// (short final range, or below the top) takes the ordinary lane solver; descents and answers
// follow per query.

// The final range's level and bounds of [l, e) as vl_start finds them, and, when it lies on the
// top level and is answered from the table, that answer's two entries and the guards, loaded
// but not yet compared.
struct VlTop {
  uint32_t K, lk, ek, shK;
  bool table, lt, rt;
  uint2 x, y;
  float gL, gR;
};

template <int RL, class LT>
__device__ __forceinline__ VlTop vl_top_issue(uint32_t l, uint32_t e, const HierView& H, const LT& T) {
  VlTop r;
  const uint32_t top = H.num_levels - 1;
  const VlLevel lv = vl_level<RL>(l, e, top);
  const uint32_t K = lv.K, shK = lv.shK, lk = lv.lk, ek = lv.ek;
  r.K = K;
  r.lk = lk;
  r.ek = ek;
  r.shK = shK;
  r.table = K == top && H.tt != nullptr && ek - lk > dev::kTopScan;
  r.lt = l < (lk << shK);
  r.rt = (ek << shK) < e;
  r.x = make_uint2(dev::f2u(RMQ_INF), 0u);
  r.y = r.x;
  r.gL = RMQ_INF;
  r.gR = RMQ_INF;
  if (r.table) {
    const float* vK = K == 0 ? H.A : H.lv + T.offz(K);
    const uint32_t j = 31u - static_cast<uint32_t>(__clz(ek - lk));
    const uint2* row = H.tt + static_cast<size_t>(j - 1u) * H.tstride;
    r.x = row[lk];
    r.y = row[ek - (1u << j)];
    if (r.lt) r.gL = vK[lk - 1u];
    if (r.rt) r.gR = vK[ek];
  }
  return r;
}

__device__ __forceinline__ VlLevel vl_level_of(const VlTop& r) { return VlLevel{r.K, r.shK, r.lk, r.ek}; }

// The lane solver's state after its final range, from vl_top_issue's loads (table case).
__device__ __forceinline__ VlState vl_top_finish(const VlTop& r, uint32_t l, uint32_t e, uint32_t RL) {
  const bool right = dev::u2f(r.y.x) < dev::u2f(r.x.x);
  const float bv = dev::u2f(right ? r.y.x : r.x.x);
  const uint32_t bpay = dev::kDirect | (right ? r.y.y : r.x.y);
  const int K = static_cast<int>(r.shK / RL);
  return VlState{l, e, bv, bpay, K - 1, r.lt && r.gL <= bv, r.rt && r.gR < bv};
}

template <int RL, bool IDX, bool W0 = false, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_vl2(const uint2* __restrict__ Q, uint32_t q,
                                                      float* __restrict__ out_v,
                                                      uint32_t* __restrict__ out_i, HierView H) {
  __shared__ uint32_t s_meta[kMetaWords];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LevelsLdsT<W0, PS> T{s_meta};
  const uint32_t zeros = *H.zeros;
  const uint64_t q0 = 2 * (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x);
  if (q0 >= q) return;
  const bool two = q0 + 1 < q;
  uint2 pa, pb;
  if (two) {  // Q is 8-byte aligned; q0 is even, so the two pairs are one 16-byte load when Q is
    const uint4 v = (reinterpret_cast<uintptr_t>(Q) & 15u) == 0
                        ? *reinterpret_cast<const uint4*>(Q + q0)
                        : make_uint4(Q[q0].x, Q[q0].y, Q[q0 + 1].x, Q[q0 + 1].y);
    pa = make_uint2(v.x, v.y);
    pb = make_uint2(v.z, v.w);
  } else {
    pa = Q[q0];
    pb = pa;
  }
  const bool da = emit_nan_at_l<IDX>(zeros, H.A, pa.x, q0, out_v, out_i);
  const bool db = !two || emit_nan_at_l<IDX>(zeros, H.A, pb.x, q0 + 1, out_v, out_i);
  // both queries' table reads in flight before either is used
  const VlTop ta = vl_top_issue<RL>(pa.x, pa.y + 1, H, T);
  const VlTop tb = vl_top_issue<RL>(pb.x, pb.y + 1, H, T);
  VlState sa = ta.table ? vl_top_finish(ta, pa.x, pa.y + 1, RL)
                        : vl_start_at<RL, IDX, true, LevelsLdsT<W0, PS>>(pa.x, pa.y + 1, vl_level_of(ta), H, T);
  VlState sb = tb.table ? vl_top_finish(tb, pb.x, pb.y + 1, RL)
                        : vl_start_at<RL, IDX, true, LevelsLdsT<W0, PS>>(pb.x, pb.y + 1, vl_level_of(tb), H, T);
  while (sa.la || sa.ra) vl_step<RL, IDX>(sa, H, T);
  while (sb.la || sb.ra) vl_step<RL, IDX>(sb, H, T);
  if (!da) vl_emit<RL, IDX>(sa, q0, H, T, zeros, out_v, out_i);
  if (!db) vl_emit<RL, IDX>(sb, q0 + 1, H, T, zeros, out_v, out_i);
}

// Two queries per lane with their descents packed (RMQ_BUILD_PAIRED | RMQ_BUILD_PACK_DESCENTS):
// both queries' final ranges (their table reads in flight together); one level of the tails,
// for both queries, where more than kInlineMin lanes of the wavefront still descend; then the
// queries still descending are parked in LDS (both queries of a lane, wavefront by wavefront)
// and after the workgroup's barrier the first threads walk them. With uniform lengths about one
// query in a hundred opens a tail after its table lookup, so without the parking most
// wavefronts of 128 queries would walk several levels for one or two lanes.
constexpr uint32_t kInlineMin = 8;
template <int RL, bool IDX, bool W0 = false, bool PS = false>
__global__ __launch_bounds__(kBlock) void k_query_vl2c(const uint2* __restrict__ Q, uint32_t q,
                                                       float* __restrict__ out_v,
                                                       uint32_t* __restrict__ out_i, HierView H) {
  using LT = LevelsLdsT<W0, PS>;
  constexpr uint32_t kWaves = kBlock / 64;
  __shared__ uint32_t s_meta[kMetaWords];
  __shared__ uint4 s_st[2 * kBlock];   // l, e, best value bits, best payload
  __shared__ uint2 s_aux[2 * kBlock];  // next level | la << 8 | ra << 9, query index
  __shared__ uint32_t s_cnt[kWaves];
  if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  __syncthreads();
  const LT T{s_meta};
  const uint32_t zeros = *H.zeros;
  const uint32_t w = threadIdx.x >> 6;
  const uint64_t q0 = 2 * (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x);
  bool need_a = false, need_b = false;
  VlState sa{}, sb{};
  if (q0 < q) {
    const bool two = q0 + 1 < q;
    uint2 pa, pb;
    if (two) {
      const uint4 v = (reinterpret_cast<uintptr_t>(Q) & 15u) == 0
                          ? *reinterpret_cast<const uint4*>(Q + q0)
                          : make_uint4(Q[q0].x, Q[q0].y, Q[q0 + 1].x, Q[q0 + 1].y);
      pa = make_uint2(v.x, v.y);
      pb = make_uint2(v.z, v.w);
    } else {
      pa = Q[q0];
      pb = pa;
    }
    const bool da = emit_nan_at_l<IDX>(zeros, H.A, pa.x, q0, out_v, out_i);
    const bool db = !two || emit_nan_at_l<IDX>(zeros, H.A, pb.x, q0 + 1, out_v, out_i);
    const VlTop ta = vl_top_issue<RL>(pa.x, pa.y + 1, H, T);
    const VlTop tb = vl_top_issue<RL>(pb.x, pb.y + 1, H, T);
    sa = ta.table ? vl_top_finish(ta, pa.x, pa.y + 1, RL)
                  : vl_start_at<RL, IDX, true, LT>(pa.x, pa.y + 1, vl_level_of(ta), H, T);
    sb = tb.table ? vl_top_finish(tb, pb.x, pb.y + 1, RL)
                  : vl_start_at<RL, IDX, true, LT>(pb.x, pb.y + 1, vl_level_of(tb), H, T);
    // one level inline where many lanes of the wavefront descend (lognormal lengths), none where
    // a few do (uniform lengths: their walks are all parked)
    if (__popcll(__ballot(sa.la || sa.ra)) > kInlineMin && (sa.la || sa.ra)) vl_step<RL, IDX>(sa, H, T);
    if (__popcll(__ballot(sb.la || sb.ra)) > kInlineMin && (sb.la || sb.ra)) vl_step<RL, IDX>(sb, H, T);
    need_a = !da && (sa.la || sa.ra);
    need_b = !db && (sb.la || sb.ra);
    if (!da && !need_a) vl_emit<RL, IDX>(sa, q0, H, T, zeros, out_v, out_i);
    if (!db && !need_b) vl_emit<RL, IDX>(sb, q0 + 1, H, T, zeros, out_v, out_i);
  }
  // wave w's parked queries take slots w * 128 + (parked queries of its lower lanes), a before b
  const uint64_t ma = __ballot(need_a), mb = __ballot(need_b);
  const uint32_t below = __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(ma >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(ma), 0u)) +
                         __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mb >> 32),
                                                   __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mb), 0u));
  const uint32_t slot = w * 128u + below;
  if (need_a) {
    s_st[slot] = make_uint4(sa.l, sa.e, dev::f2u(sa.bv), sa.bpay);
    s_aux[slot] = make_uint2(static_cast<uint32_t>(sa.k) | (sa.la ? 0x100u : 0u) | (sa.ra ? 0x200u : 0u),
                             static_cast<uint32_t>(q0));
  }
  if (need_b) {
    const uint32_t sl = slot + (need_a ? 1u : 0u);
    s_st[sl] = make_uint4(sb.l, sb.e, dev::f2u(sb.bv), sb.bpay);
    s_aux[sl] = make_uint2(static_cast<uint32_t>(sb.k) | (sb.la ? 0x100u : 0u) | (sb.ra ? 0x200u : 0u),
                           static_cast<uint32_t>(q0 + 1));
  }
  if ((threadIdx.x & 63u) == 0u) s_cnt[w] = static_cast<uint32_t>(__popcll(ma) + __popcll(mb));
  __syncthreads();
  uint32_t c[kWaves], total = 0;
#pragma unroll
  for (uint32_t v = 0; v < kWaves; ++v) {
    c[v] = s_cnt[v];
    total += c[v];
  }
  for (uint32_t j = threadIdx.x; j < total; j += kBlock) {
    uint32_t r = j, sl = 2 * kBlock;
#pragma unroll
    for (uint32_t v = 0; v < kWaves; ++v) {
      if (sl == 2 * kBlock) {
        if (r < c[v]) sl = v * 128u + r;
        else r -= c[v];
      }
    }
    const uint4 a = s_st[sl];
    const uint2 b = s_aux[sl];
    VlState ps{a.x, a.y, dev::u2f(a.z), a.w, static_cast<int>(b.x & 0xFFu), (b.x & 0x100u) != 0u,
               (b.x & 0x200u) != 0u};
    while (ps.la || ps.ra) vl_step<RL, IDX>(ps, H, T);
    vl_emit<RL, IDX>(ps, b.y, H, T, zeros, out_v, out_i);
  }
}

// ---------------------------------------------------------------------------------------
// Two queries per lane with the top level and its sparse table in LDS (the uniform-length
// form). A query answered from the table reads its pair, then two table entries and its guards,
// and writes its answer: with the table in L2 the middle step is a dependent L2 round trip per
// query. Here every workgroup (kTopBlock lanes, persistent) first copies the top level's values
// and the table rows it reads (rows j >= kLdsTopRow0: final ranges of at least 16 top entries;
// shorter ones take the ordinary solver, which reads the table in HBM above 8 entries) into
// LDS; those reads are then LDS reads. With uniform lengths about one query in a hundred opens
// a tail after its table lookup, and a wavefront would walk it for one or two lanes: instead
// the queries still descending are parked in the wavefront's own LDS queue (kParkCap entries;
// no barrier), and walked kParkCap at a time on all lanes when it would overflow and at the
// end. Top levels of 16..kLdsTopMax entries.
constexpr uint32_t kTopBlock = 1024;
constexpr uint32_t kLdsTopMax = 1024;
constexpr uint32_t kLdsTopRow0 = 4;
constexpr uint32_t kParkCap = 64;
constexpr uint32_t kUnstarted = 0x400u;  // a parked query whose solve has not started

// LDS bytes of that image for a top level of m entries (rows = floor(log2 m) table rows), with
// the wavefronts' park queues after it.
__host__ __device__ constexpr uint32_t lds_top_image(uint32_t m, uint32_t rows) {
  return ((4u * m + 15u) & ~15u) + (rows + 1u - kLdsTopRow0) * m * 8u;
}
__host__ __device__ constexpr uint32_t lds_top_bytes(uint32_t m, uint32_t rows) {
  return lds_top_image(m, rows) + (kTopBlock / 64u) * kParkCap * 16u;
}

// At most 80 scalar registers: above that (91 measured in index mode) a CU holds only one
// workgroup of 1024 lanes instead of two, and the kernel ran 118 instead of 162 Gq/s (uniform
// lengths, n = q = 2^26); the compiler's own occupancy estimate does not show it.
template <int RL, bool IDX, bool W0 = false, bool PS = false>
__global__ __launch_bounds__(kTopBlock) __attribute__((amdgpu_num_sgpr(80))) void k_query_vl2t(
    const uint2* __restrict__ Q, uint32_t q, float* __restrict__ out_v, uint32_t* __restrict__ out_i, HierView H) {
  using LT = LevelsLdsT<W0, PS>;
  extern __shared__ __align__(16) uint32_t s_dyn[];
  __shared__ uint32_t s_meta[kMetaWords];
  const uint32_t top = H.num_levels - 1;
  const uint32_t m = H.size[top];
  const uint32_t rows = 31u - static_cast<uint32_t>(__clz(m));
  float* s_topv = reinterpret_cast<float*>(s_dyn);  // m values
  uint2* s_tt = reinterpret_cast<uint2*>(s_dyn + ((m + 3u) & ~3u));  // rows kLdsTopRow0.. of the table
  const uint32_t lane = threadIdx.x & 63u;
  // this wavefront's park queue: {query index, best value bits, best payload, next level | la | ra}
  uint4* s_park = reinterpret_cast<uint4*>(reinterpret_cast<uint8_t*>(s_dyn) + lds_top_image(m, rows)) +
                  (threadIdx.x >> 6) * kParkCap;
  {
    const float* tv = H.lv + H.off[top];
    for (uint32_t i = threadIdx.x; i < m; i += kTopBlock) s_topv[i] = tv[i];
    const uint32_t cnt = (rows + 1u - kLdsTopRow0) * m;
    for (uint32_t x = threadIdx.x; x < cnt; x += kTopBlock) {
      const uint32_t r = x / m, i = x - r * m;
      s_tt[x] = H.tt[static_cast<size_t>(r + kLdsTopRow0 - 1u) * H.tstride + i];
    }
    if (threadIdx.x < kMetaWords) s_meta[threadIdx.x] = lds_meta(H, threadIdx.x);
  }
  __syncthreads();
  const LT T{s_meta};
  const uint32_t zeros = *H.zeros;
  // the final range of [l, e) when it lies on the top level and spans at least 2^kLdsTopRow0
  // entries: its table answer and guards from LDS (false: the ordinary solver)
  auto from_table = [&](uint32_t l, uint32_t e, VlState& st) {
    const uint32_t K = min((31u - static_cast<uint32_t>(__clz(e - l))) / RL, top);
    const uint32_t shK = RL * K;
    const uint32_t lk = (l >> shK) + ((l & ((1u << shK) - 1u)) != 0u);
    const uint32_t ek = e >> shK;
    if (K != top || lk >= ek || ek - lk < (1u << kLdsTopRow0)) return false;
    const uint32_t j = 31u - static_cast<uint32_t>(__clz(ek - lk));
    const uint2* row = s_tt + (j - kLdsTopRow0) * m;
    const uint2 a = row[lk], b = row[ek - (1u << j)];
    const bool right = dev::u2f(b.x) < dev::u2f(a.x);
    const float bv = dev::u2f(right ? b.x : a.x);
    const bool lt = l < (lk << shK), rt = (ek << shK) < e;
    const float gL = lt ? s_topv[lk - 1u] : RMQ_INF;
    const float gR = rt ? s_topv[ek] : RMQ_INF;
    st = VlState{l, e, bv, dev::kDirect | (right ? b.y : a.y), static_cast<int>(K) - 1, lt && gL <= bv,
                 rt && gR < bv};
    return true;
  };
  // Both modes track the winner's position: its value is then the bit pattern of A[leftmost
  // argmin] (the table's entries hold it, and the solver keeps the winning element's own bits),
  // so value mode needs no second solve for a zero of unknown sign.
  auto emit = [&](const VlState& st, uint64_t qi) {
    if (out_v) out_v[qi] = st.bv;
    if (IDX && out_i) out_i[qi] = pay_index<RL>(st.bpay, H, T);
  };
  // the park queue (its count is the same in every lane): walked on the first `parked` lanes
  uint32_t parked = 0;
  auto walk_parked = [&]() {
    __builtin_amdgcn_wave_barrier();
    if (lane < parked) {
      const uint4 x = s_park[lane];
      const uint2 p = Q[x.x];
      VlState st = (x.w & kUnstarted) ? vl_start<RL, true, true, LT>(p.x, p.y + 1, H, T)
                                      : VlState{p.x, p.y + 1, dev::u2f(x.y), x.z, static_cast<int>(x.w & 0xFFu),
                                                (x.w & 0x100u) != 0u, (x.w & 0x200u) != 0u};
      while (st.la || st.ra) vl_step<RL, true>(st, H, T);
      emit(st, x.x);
    }
    __builtin_amdgcn_wave_barrier();
    parked = 0;
  };
  // a query that still descends is parked (the queue walked first if it would overflow);
  // true if it was
  auto park = [&](bool open, const VlState& st, uint32_t qi, bool unstarted) {
    const uint64_t mk = __ballot(open);
    const uint32_t c = static_cast<uint32_t>(__popcll(mk));
    if (c == 0u) return false;
    if (parked + c > kParkCap) walk_parked();
    if (open) {
      const uint32_t at = parked + __builtin_amdgcn_mbcnt_hi(static_cast<uint32_t>(mk >> 32),
                                                             __builtin_amdgcn_mbcnt_lo(static_cast<uint32_t>(mk), 0u));
      s_park[at] = make_uint4(qi, dev::f2u(st.bv), st.bpay,
                              unstarted ? kUnstarted
                                        : static_cast<uint32_t>(st.k) | (st.la ? 0x100u : 0u) | (st.ra ? 0x200u : 0u));
    }
    parked += c;
    return open;
  };
  const uint64_t step = 2ull * kTopBlock * gridDim.x;
  // the wavefront stays in the loop until its last pair (its queue is walked on all its lanes)
  for (uint64_t w0 = 2ull * (static_cast<uint64_t>(blockIdx.x) * kTopBlock + (threadIdx.x & ~63u)); w0 < q;
       w0 += step) {
    const uint64_t q0 = w0 + 2u * lane;
    const bool one = q0 < q, two = q0 + 1 < q;
    uint2 pa = make_uint2(0u, 0u), pb = pa;
    if (two) {
      const uint4 v = (reinterpret_cast<uintptr_t>(Q) & 15u) == 0
                          ? *reinterpret_cast<const uint4*>(Q + q0)
                          : make_uint4(Q[q0].x, Q[q0].y, Q[q0 + 1].x, Q[q0 + 1].y);
      pa = make_uint2(v.x, v.y);
      pb = make_uint2(v.z, v.w);
    } else if (one) {
      pa = Q[q0];
      pb = pa;
    }
    const bool da = !one || emit_nan_at_l<IDX>(zeros, H.A, pa.x, q0, out_v, out_i);
    const bool db = !two || emit_nan_at_l<IDX>(zeros, H.A, pb.x, q0 + 1, out_v, out_i);
    // a query not answered from the table is parked unstarted (its whole solve then runs on a
    // full wavefront), one that still descends is parked mid-walk
    VlState sa{}, sb{};
    const bool ta = !da && from_table(pa.x, pa.y + 1, sa);
    const bool tb = !db && from_table(pb.x, pb.y + 1, sb);
    const bool pka = park(!da && (!ta || sa.la || sa.ra), sa, static_cast<uint32_t>(q0), !ta);
    const bool pkb = park(!db && (!tb || sb.la || sb.ra), sb, static_cast<uint32_t>(q0 + 1), !tb);
    if (!da && !pka) emit(sa, q0);
    if (!db && !pkb) emit(sb, q0 + 1);
  }
  if (parked) walk_parked();
}

}  // namespace rmq
