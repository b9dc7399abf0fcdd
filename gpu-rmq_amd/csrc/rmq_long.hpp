// rmq_long.hpp — the 64-bit form (the reference's IS_LONG build: 64-bit positions and {l, r}
// pairs, src/main.cu:21-27, src/tools.h:27-31) for arrays of 2^31 elements and more.
//
// Such an array is held as chunks of 2^c elements (c = 30 unless the caller asks for less),
// each an ordinary 32-bit hierarchy, plus every chunk's minimum and its leftmost index. A query
// inside one chunk is that chunk's query; a query across chunks is the suffix query of its first
// chunk, the minima of the chunks it covers whole, and the prefix query of its last chunk,
// combined left to right with strict `<` so the leftmost minimum wins. Every part is answered
// by the lane solver of the chunks' reduction factor, with index tracking, so values carry the
// bit pattern of A[leftmost argmin] without a zero-sign rule.
#pragma once

#include "rmq_kernels.hpp"

namespace rmq {

template <int RL, class LT>
__device__ __forceinline__ GroupAnswer solve_chunk(const HierView& H, const LT& T, uint32_t l, uint32_t e) {
  return solve_vl_pruned<RL, true>(l, e, H, T);
}

// Chunks whose level tables a query kernel copies into LDS (more fall back to reading them from
// the chunks' views).
constexpr uint32_t kLdsChunks = 64;

// The minimum of every whole chunk: one lane per chunk.
template <int RL>
__global__ __launch_bounds__(kBlock) void k_chunk_min(const HierView* __restrict__ views, uint32_t chunks,
                                                      uint32_t clog, float* __restrict__ cv,
                                                      uint64_t* __restrict__ ci) {
  const uint32_t c = blockIdx.x * kBlock + threadIdx.x;
  if (c >= chunks) return;
  const HierView& H = views[c];
  const GroupAnswer a = solve_chunk<RL>(H, LevelsOf{H}, 0, H.size[0]);
  cv[c] = a.v;
  ci[c] = (static_cast<uint64_t>(c) << clog) + a.idx;
}

// One lane per query; Q holds {l, r} pairs of 64-bit positions. LDS_META: all `chunks` level
// tables (at most kLdsChunks) are copied into LDS first.
template <int RL, bool LDS_META>
__global__ __launch_bounds__(kBlock) void k_query64(const ulonglong2* __restrict__ Q, uint64_t q,
                                                    float* __restrict__ out_v, uint64_t* __restrict__ out_i,
                                                    const HierView* __restrict__ views, uint32_t clog,
                                                    uint32_t chunks, const float* __restrict__ cv,
                                                    const uint64_t* __restrict__ ci) {
  __shared__ uint32_t s_meta[LDS_META ? kLdsChunks * kMetaWords : 1];
  if constexpr (LDS_META) {
    for (uint32_t i = threadIdx.x; i < chunks * kMetaWords; i += kBlock)
      s_meta[i] = lds_meta(views[i / kMetaWords], i % kMetaWords);
    __syncthreads();
  }
  auto table = [&](uint32_t c) {
    if constexpr (LDS_META) return LevelsLds{s_meta + c * kMetaWords};
    else return LevelsOf{views[c]};
  };
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= q) return;
  const ulonglong2 p = Q[qi];
  const uint64_t l = p.x, e = p.y + 1;
  const uint32_t cl = static_cast<uint32_t>(l >> clog), cr = static_cast<uint32_t>(p.y >> clog);
  const uint64_t bl = static_cast<uint64_t>(cl) << clog;
  // every query solves its part in its first chunk (all of it, or the chunk's suffix from l); a
  // query across chunks then adds the whole chunks' minima and its last chunk's prefix, so a
  // wavefront that mixes both kinds runs the solver twice, not three times
  const HierView& HL = views[cl];
  const uint32_t ea = cl == cr ? static_cast<uint32_t>(e - bl) : HL.size[0];
  const GroupAnswer a = solve_chunk<RL>(HL, table(cl), static_cast<uint32_t>(l - bl), ea);
  float bv = a.v;
  uint64_t bi = bl + a.idx;
  // CPU BASE's NaN rule, for the query's own first element only: the chunks' minima and the
  // last chunk's prefix are answered without it (a NaN there never wins)
  if (dev::nan_at_l(*HL.zeros, HL.A, l - bl, bv)) {
    bi = l;
  } else if (cl != cr) {
    for (uint32_t c = cl + 1; c < cr; ++c) {  // chunks covered whole, left to right
      const float v = cv[c];
      if (v < bv) {
        bv = v;
        bi = ci[c];
      }
    }
    const uint64_t br = static_cast<uint64_t>(cr) << clog;
    const GroupAnswer b = solve_chunk<RL>(views[cr], table(cr), 0, static_cast<uint32_t>(e - br));
    if (b.v < bv) {
      bv = b.v;
      bi = br + b.idx;
    }
  }
  if (out_v) out_v[qi] = bv;
  if (out_i) out_i[qi] = bi;
}

// GPU BASE with 64-bit positions: one lane per query scans [l, r] with strict `<`.
__global__ __launch_bounds__(kBlock) void k_gpu_base64(const float* __restrict__ A,
                                                       const ulonglong2* __restrict__ Q, uint64_t q,
                                                       float* __restrict__ out_v,
                                                       uint64_t* __restrict__ out_i) {
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= q) return;
  const ulonglong2 p = Q[qi];
  float m = A[p.x];
  uint64_t mi = p.x;
  for (uint64_t i = p.x + 1; i <= p.y; ++i) {
    const float x = A[i];
    if (x < m) {
      m = x;
      mi = i;
    }
  }
  if (out_v) out_v[qi] = m;
  if (out_i) out_i[qi] = mi;
}

}  // namespace rmq
