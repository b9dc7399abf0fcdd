// rmq_kernels.hpp — gfx950 kernels of the hierarchical RMQ: level build, the three query
// variants (algs 16, 24, 20) and GPU BASE (alg 2/102).
//
// Data layout in HBM: level 0 is the caller's float array (borrowed). Levels 1.. live in one
// value buffer and one index buffer at the reference's offsets (multiples of 32 elements,
// src/xxx.h:238), so every level starts 128-byte aligned and every R-block of R >= 4
// elements is whole 16-byte vectors.
#pragma once

#include "rmq_device.hpp"

namespace rmq {

constexpr int kBlock = 256;  // 4 waves per workgroup

// ---------------------------------------------------------------------------------------
// Build: level k+1 entry j = min of level k over [j*R, min(j*R+R, size_k)) with the leftmost
// level-0 index of that minimum (reference: xxx_build_step, src/xxx.h:91-120, which keeps
// the value only). W = R/4 lanes share one block, each lane owns 16-byte vectors, so one
// wave instruction streams 1 KiB of contiguous level-k data; the W lanes meet in a DPP
// all-reduce. Grid-stride over output entries.
template <int RL, bool SRC_IDX>
__global__ __launch_bounds__(kBlock) void k_build_level(const float* __restrict__ sv,
                                                        const uint32_t* __restrict__ si,
                                                        uint32_t ssize, float* __restrict__ dv,
                                                        uint32_t* __restrict__ di,
                                                        uint32_t dsize) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;                 // 16-byte vectors per block
  constexpr uint32_t W = V < 64 ? V : 64;       // lanes per block
  constexpr uint32_t PER = V / W;               // vectors per lane
  const uint32_t gl = threadIdx.x & (W - 1);
  const uint64_t grp = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / W;
  const uint64_t ngrp = static_cast<uint64_t>(gridDim.x) * kBlock / W;
  for (uint64_t j = grp; j < dsize; j += ngrp) {
    float bv = RMQ_INF;
    uint32_t bi = dev::kNoIdx;
    const uint32_t base = static_cast<uint32_t>(j) * R;
#pragma unroll
    for (uint32_t p = 0; p < PER; ++p) {
      const uint32_t pos = base + 4 * (p * W + gl);
      if (pos < ssize) {
        const float4 x = dev::load4(sv, pos, ssize);
        dev::fold4_idx(bv, bi, x, SRC_IDX ? si : nullptr, pos, ssize, pos, ssize);
      }
    }
    dev::group_argmin<W>(bv, bi);
    if (gl == 0) {
      dv[j] = bv;
      di[j] = bi;
    }
  }
}

// ---------------------------------------------------------------------------------------
// Query traversal shared by algs 16 and 24 (src/xxx.h:158-187 and 493-521): r exclusive;
// while the remaining range at this level is longer than 2R, fold the two partial blocks
// [l, ceil_R(l)) and [floor_R(r), r) and climb; finally scan [l, r) at the level reached.
// A group of G lanes works on one query: each lane owns 16-byte vectors of a block, so a
// 32-float block is one vector load per lane at G = 8. The level loop runs with a
// wave-uniform counter (metadata stays in scalar registers); groups that stopped climbing
// sit it out.
template <int G, int RL, bool IDX>
__device__ __forceinline__ void solve_cl(uint32_t l, uint32_t e, uint32_t gl, const HierView& H,
                                         float& out_v, uint32_t& out_i) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;   // vectors per block
  float bv = RMQ_INF;
  uint32_t bi = dev::kNoIdx;
  const float* cur_v = H.A;
  const uint32_t* cur_i = nullptr;
  uint32_t cur_size = H.size[0];
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
          if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p, cur_size, l, nl);
          else bv = dev::fold4(bv, x, p, l, nl);
        }
        const uint32_t p2 = pr + 4 * s;
        if (p2 < e) {
          const float4 x = dev::load4(cur_v, p2, cur_size);
          if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p2, cur_size, pr, e);
          else bv = dev::fold4(bv, x, p2, pr, e);
        }
      }
      l = nl >> RL;
      e = pr >> RL;
      cur_v = H.lv + H.off[lev + 1];
      cur_i = H.li + H.off[lev + 1];
      cur_size = H.size[lev + 1];
    } else {
      climbing = false;
    }
  }
  // final scan of [l, e) at the level reached, 4*G elements per step
  for (uint32_t p = (l & ~(4 * G - 1)) + 4 * gl; p < e; p += 4 * G) {
    if (p + 4 > l) {
      const float4 x = dev::load4(cur_v, p, cur_size);
      if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p, cur_size, l, e);
      else bv = dev::fold4(bv, x, p, l, e);
    }
  }
  if constexpr (IDX) {
    dev::group_argmin<G>(bv, bi);
  } else {
    bv = dev::group_min<G>(bv);
  }
  out_v = bv;
  out_i = bi;
}

// Value-only answers must carry the bit pattern of A[leftmost argmin]. A plain minimum can
// only differ from that when the minimum is zero (+0 and -0 compare equal); those queries
// are answered again with index tracking.
template <int G, int RL>
__device__ __forceinline__ float value_exact_cl(uint32_t l, uint32_t e, uint32_t gl,
                                                const HierView& H) {
  float v;
  uint32_t i;
  solve_cl<G, RL, false>(l, e, gl, H, v, i);
  if (v == 0.0f) solve_cl<G, RL, true>(l, e, gl, H, v, i);
  return v;
}

// Alg 16 (xxx_query, src/xxx.h:124-200): every lane owns one query; a G-lane group serves
// its members' queries one after another (the reference's ballot work queue, here a fixed
// round over the group's lanes with a ds_bpermute broadcast). Queries are read and answers
// written with one coalesced access per wave.
template <int G, int RL, bool IDX>
__global__ __launch_bounds__(kBlock) void k_query_xxx(const uint2* __restrict__ Q, uint32_t q,
                                                      float* __restrict__ out_v,
                                                      uint32_t* __restrict__ out_i, HierView H) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gl = lane & (G - 1);
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  const bool valid = qi < q;
  uint32_t l = 0, e = 0;
  if (valid) {
    const uint2 p = Q[qi];
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
        solve_cl<G, RL, true>(lk, ek, gl, H, v, i);
      } else {
        v = value_exact_cl<G, RL>(lk, ek, gl, H);
      }
      if (gl == static_cast<uint32_t>(k)) {
        my_v = v;
        my_i = i;
      }
    }
  }
  if (valid) {
    if (out_v) out_v[qi] = my_v;
    if (IDX && out_i) out_i[qi] = my_i;
  }
}

// Alg 24 (xxx_query_multi_load, src/xxx.h:470-528): one G-lane group per query, every lane of
// the group reads the same query; grid-stride over queries in units of groups (the
// reference strides by the thread count, Appendix-A quirk not kept).
template <int G, int RL, bool IDX>
__global__ __launch_bounds__(kBlock) void k_query_multi(const uint2* __restrict__ Q, uint32_t q,
                                                        float* __restrict__ out_v,
                                                        uint32_t* __restrict__ out_i,
                                                        HierView H) {
  const uint32_t gl = threadIdx.x & (G - 1);
  const uint64_t grp = (static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x) / G;
  const uint64_t ngrp = static_cast<uint64_t>(gridDim.x) * kBlock / G;
  for (uint64_t qi = grp; qi < q; qi += ngrp) {
    const uint2 p = Q[qi];
    float v;
    uint32_t i = dev::kNoIdx;
    if constexpr (IDX) {
      solve_cl<G, RL, true>(p.x, p.y + 1, gl, H, v, i);
    } else {
      v = value_exact_cl<G, RL>(p.x, p.y + 1, gl, H);
    }
    if (gl == 0) {
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
__device__ __forceinline__ void solve_vl(uint32_t l, uint32_t e, const HierView& H, float& out_v,
                                         uint32_t& out_i) {
  constexpr uint32_t R = 1u << RL;
  constexpr uint32_t V = R / 4;
  float bv = RMQ_INF;
  uint32_t bi = dev::kNoIdx;
  const float* cur_v = H.A;
  const uint32_t* cur_i = nullptr;
  uint32_t cur_size = H.size[0];
  bool climbing = true;
  for (uint32_t lev = 0; lev + 1 < H.num_levels; ++lev) {
    if (climbing && e - l > R) {
      const uint32_t nl = (l + R - 1) & ~(R - 1);
      const uint32_t pr = e & ~(R - 1);
      if (pr >= nl) {
        const uint32_t lb = l & ~(R - 1);
#pragma unroll
        for (uint32_t s = 0; s < V; ++s) {
          const uint32_t p = lb + 4 * s;
          if (p + 4 > l && p < nl) {
            const float4 x = dev::load4(cur_v, p, cur_size);
            if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p, cur_size, l, nl);
            else bv = dev::fold4(bv, x, p, l, nl);
          }
        }
#pragma unroll
        for (uint32_t s = 0; s < V; ++s) {
          const uint32_t p = pr + 4 * s;
          if (p < e) {
            const float4 x = dev::load4(cur_v, p, cur_size);
            if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p, cur_size, pr, e);
            else bv = dev::fold4(bv, x, p, pr, e);
          }
        }
        l = nl >> RL;
        e = pr >> RL;
        cur_v = H.lv + H.off[lev + 1];
        cur_i = H.li + H.off[lev + 1];
        cur_size = H.size[lev + 1];
      } else {
        climbing = false;
      }
    } else {
      climbing = false;
    }
  }
  for (uint32_t p = l & ~3u; p < e; p += 4) {
    const float4 x = dev::load4(cur_v, p, cur_size);
    if constexpr (IDX) dev::fold4_idx(bv, bi, x, cur_i, p, cur_size, l, e);
    else bv = dev::fold4(bv, x, p, l, e);
  }
  out_v = bv;
  out_i = bi;
}

template <int RL, bool IDX>
__global__ __launch_bounds__(kBlock) void k_query_vl(const uint2* __restrict__ Q, uint32_t q,
                                                     float* __restrict__ out_v,
                                                     uint32_t* __restrict__ out_i, HierView H) {
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= q) return;
  const uint2 p = Q[qi];
  float v;
  uint32_t i = dev::kNoIdx;
  if constexpr (IDX) {
    solve_vl<RL, true>(p.x, p.y + 1, H, v, i);
  } else {
    solve_vl<RL, false>(p.x, p.y + 1, H, v, i);
    if (v == 0.0f) solve_vl<RL, true>(p.x, p.y + 1, H, v, i);
  }
  if (out_v) out_v[qi] = v;
  if (IDX && out_i) out_i[qi] = i;
}

// ---------------------------------------------------------------------------------------
// GPU BASE (kernel_rmq_basic / _idx / value+index, src/cuda_methods.cuh:4-68): one lane per
// query, serial strict-`<` scan from l to r. The reference's --check oracle.
__global__ __launch_bounds__(kBlock) void k_gpu_base(const float* __restrict__ A,
                                                     const uint2* __restrict__ Q, uint32_t q,
                                                     float* __restrict__ out_v,
                                                     uint32_t* __restrict__ out_i) {
  const uint64_t qi = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  if (qi >= q) return;
  const uint2 p = Q[qi];
  float m = A[p.x];
  uint32_t mi = p.x;
  for (uint32_t i = p.x; i <= p.y; ++i) {
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
