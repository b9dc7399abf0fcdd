// rmq_multi.hpp — multi-array query machinery shared by the batched form (rmq_batch.hip) and
// the serving loop (rmq_serve.hip): a table of per-array HierViews in global memory, 256-query
// tiles, and one persistent kernel that pulls tiles from per-XCD work queues.
//
// XCD-aware scheduling: workgroups are dispatched round-robin over the 8 XCDs of an MI355X,
// so workgroup b runs on XCD b % 8. Arrays are assigned to 8 queues and workgroups pull from
// the queue of their own XCD first, so an array's upper levels are cached in one XCD's L2; an
// idle workgroup then steals from the other queues. Correctness does not depend on the
// dispatch order: every tile is taken exactly once through an atomic counter.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/rmq_batch.h"
#include "rmq_kernels.hpp"

namespace rmq {

constexpr uint32_t kMultiRLog = 5;       // R = 32 for every array of a batch
constexpr uint32_t kMultiR = 1u << kMultiRLog;
constexpr uint32_t kQueues = 8;          // XCDs of an MI355X
constexpr uint32_t kTileQueries = kBlock;

struct QueryTile {
  uint32_t array;
  uint32_t count;  // queries in the tile (<= 256)
  uint64_t q0;     // first query index (in the query / output buffers of the launch)
};

template <int N>
__global__ void k_reset_counters(uint32_t* __restrict__ c) {
  if (threadIdx.x < N) c[threadIdx.x] = 0;
}

// ---- query: persistent, tiles pulled from per-XCD queues ----------------------------------------
template <int G, bool IDX>
__device__ __forceinline__ void tile_groups(const HierView& H, const uint2* __restrict__ Q,
                                            uint64_t q0, uint32_t count, float* __restrict__ ov,
                                            uint32_t* __restrict__ oi) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gl = lane & (G - 1);
  const uint32_t local = threadIdx.x;  // one query per lane, 256 per tile
  const bool valid = local < count;
  uint32_t l = 0, e = 0;
  if (valid) {
    const uint2 p = dev::load_query(Q, q0 + local);
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
        const GroupAnswer a = solve_cl_pruned<G, kMultiRLog, true>(lk, ek, gl, H);
        v = a.v;
        i = a.idx;
      } else {
        v = value_exact<G, kMultiRLog, true>(lk, ek, gl, H);
      }
      if (gl == static_cast<uint32_t>(k)) {
        my_v = v;
        my_i = i;
      }
    }
  }
  if (valid) {
    if (ov) ov[q0 + local] = my_v;
    if (IDX && oi) oi[q0 + local] = my_i;
  }
}

template <bool IDX>
__device__ __forceinline__ void tile_lanes(const HierView& H, const uint2* __restrict__ Q,
                                           uint64_t q0, uint32_t count, float* __restrict__ ov,
                                           uint32_t* __restrict__ oi) {
  const uint32_t local = threadIdx.x;
  if (local >= count) return;
  const uint2 p = dev::load_query(Q, q0 + local);
  GroupAnswer a = solve_vl_pruned<kMultiRLog, IDX>(p.x, p.y + 1, H);
  if constexpr (!IDX) {
    if (a.v == 0.0f) a = solve_vl_pruned<kMultiRLog, true>(p.x, p.y + 1, H);
  }
  if (ov) ov[q0 + local] = a.v;
  if (IDX && oi) oi[q0 + local] = a.idx;
}

template <bool IDX>
__global__ __launch_bounds__(kBlock) void k_query_batch(
    const HierView* __restrict__ views, const uint32_t* __restrict__ variant,
    const QueryTile* __restrict__ tiles, const uint32_t* __restrict__ qstart,
    uint32_t* __restrict__ counter, const uint2* __restrict__ Q, float* __restrict__ ov,
    uint32_t* __restrict__ oi) {
  __shared__ uint32_t s_tile;
  const uint32_t home = blockIdx.x % kQueues;
  uint32_t queue = home;
  uint32_t tried = 0;
  while (tried < kQueues) {
    if (threadIdx.x == 0) {
      const uint32_t len = qstart[queue + 1] - qstart[queue];
      const uint32_t t = atomicAdd(&counter[queue], 1u);
      s_tile = t < len ? qstart[queue] + t : 0xFFFFFFFFu;
    }
    __syncthreads();
    const uint32_t tile = s_tile;
    __syncthreads();
    if (tile == 0xFFFFFFFFu) {  // this queue is drained: move to the next one
      queue = (queue + 1) % kQueues;
      ++tried;
      continue;
    }
    const QueryTile T = tiles[tile];
    const uint32_t a = __builtin_amdgcn_readfirstlane(T.array);
    const HierView& H = views[a];
    switch (variant[a]) {
      case RMQ_VARIANT_GROUP4: tile_groups<4, IDX>(H, Q, T.q0, T.count, ov, oi); break;
      case RMQ_VARIANT_GROUP16: tile_groups<16, IDX>(H, Q, T.q0, T.count, ov, oi); break;
      case RMQ_VARIANT_LANE: tile_lanes<IDX>(H, Q, T.q0, T.count, ov, oi); break;
      default: tile_groups<8, IDX>(H, Q, T.q0, T.count, ov, oi); break;
    }
  }
}

// Tiles of 256 queries per segment, grouped into kQueues queues balanced by query count.
// qoffsets[a]..qoffsets[a+1] are segment a's queries, on array ids[a] (ids == nullptr: array a);
// returns the tiles and the queue starts.
inline void make_query_tiles(uint32_t count, const uint64_t* qoffsets, std::vector<QueryTile>& tiles,
                             std::vector<uint32_t>& qstart, const uint32_t* ids = nullptr) {
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  std::sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
    return qoffsets[x + 1] - qoffsets[x] > qoffsets[y + 1] - qoffsets[y];
  });
  std::vector<uint64_t> load(kQueues, 0);
  std::vector<std::vector<uint32_t>> members(kQueues);
  for (uint32_t a : order) {
    const uint32_t qn = static_cast<uint32_t>(std::min_element(load.begin(), load.end()) - load.begin());
    load[qn] += qoffsets[a + 1] - qoffsets[a];
    members[qn].push_back(a);
  }
  tiles.clear();
  qstart.assign(kQueues + 1, 0);
  for (uint32_t qn = 0; qn < kQueues; ++qn) {
    qstart[qn] = static_cast<uint32_t>(tiles.size());
    std::sort(members[qn].begin(), members[qn].end());
    for (uint32_t a : members[qn]) {
      for (uint64_t q = qoffsets[a]; q < qoffsets[a + 1]; q += kTileQueries) {
        const uint64_t c = std::min<uint64_t>(kTileQueries, qoffsets[a + 1] - q);
        tiles.push_back(QueryTile{ids ? ids[a] : a, static_cast<uint32_t>(c), q});
      }
    }
  }
  qstart[kQueues] = static_cast<uint32_t>(tiles.size());
}

inline uint32_t persistent_grid(const void* kernel) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 1024;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, kBlock, 0) != hipSuccess || per_cu < 1)
    per_cu = 2;
  const uint32_t g = static_cast<uint32_t>(cus) * static_cast<uint32_t>(per_cu);
  return (g + kQueues - 1) / kQueues * kQueues;
}

// Counter reset + persistent query launch over the tiles of one plan.
inline void launch_multi_query(const HierView* views, const uint32_t* variant, const QueryTile* tiles,
                               const uint32_t* qstart, uint32_t* counter, const uint2* Q, float* ov,
                               uint32_t* oi, hipStream_t s) {
  k_reset_counters<kQueues><<<1, 64, 0, s>>>(counter);
  if (oi) {
    static const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(k_query_batch<true>));
    k_query_batch<true><<<grid, kBlock, 0, s>>>(views, variant, tiles, qstart, counter, Q, ov, oi);
  } else {
    static const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(k_query_batch<false>));
    k_query_batch<false><<<grid, kBlock, 0, s>>>(views, variant, tiles, qstart, counter, Q, ov, oi);
  }
}

}  // namespace rmq
