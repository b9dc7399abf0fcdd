// rmq_multi.hpp — multi-array query machinery shared by the batched form (rmq_batch.hip) and
// the serving loop (rmq_serve.hip): a table of per-array HierViews in global memory, 256-query
// tiles, and one persistent kernel whose wavefronts pull work from per-XCD queues.
//
// XCD-aware scheduling: workgroups are dispatched round-robin over the 8 XCDs of an MI355X,
// so workgroup b runs on XCD b % 8. The tiles, ordered by array, are cut into 8 contiguous
// queues of about equal query count, and the wavefronts of a workgroup work on the queue of
// their own XCD first, so an array's upper levels are cached in few XCDs' L2. A queue's first
// half is split statically among its home wavefronts (no atomics); the second half is handed
// out in 64-query units, kClaimUnits per atomic increment of the queue's counter, so that
// wavefronts that finish early balance the rest; an idle wavefront then steals from the other
// queues. Each counter sits on its own 256-byte line. Global atomics are performed past the
// XCD-private L2s, so every increment is a fabric round trip and increments of one counter
// serialize: the static half and the multi-unit claims keep them to a few thousand per launch.
// Correctness does not depend on the dispatch order: every unit is taken exactly once.
#pragma once

#include <algorithm>
#include <cstdint>
#include <numeric>
#include <vector>

#include "../../include/rmq_batch.h"
#include "rmq_kernels.hpp"

namespace rmq {

constexpr uint32_t kMultiRLog = 5;       // R = 32: group variants and RMQ_VARIANT_LANE
constexpr uint32_t kLane8RLog = 3;       // R = 8: RMQ_VARIANT_LANE8
constexpr uint32_t kLane4RLog = 2;       // R = 4: RMQ_VARIANT_LANE4
constexpr uint32_t kMultiR = 1u << kMultiRLog;

// Reduction factor (log2) and plan parameters of a variant.
__host__ __device__ inline uint32_t variant_rlog(uint32_t variant) {
  return variant == RMQ_VARIANT_LANE8 ? kLane8RLog : (variant == RMQ_VARIANT_LANE4 ? kLane4RLog : kMultiRLog);
}

inline bool variant_ok(uint32_t v) {
  return v == RMQ_VARIANT_GROUP4 || v == RMQ_VARIANT_GROUP8 || v == RMQ_VARIANT_GROUP16 ||
         v == RMQ_VARIANT_GROUP32 || v == RMQ_VARIANT_WAVE || v == RMQ_VARIANT_LANE ||
         v == RMQ_VARIANT_LANE8 || v == RMQ_VARIANT_LANE4;
}

// A variant with one lane per query (the lanes-only kernel serves launches of these alone).
__host__ __device__ inline bool variant_is_lane(uint32_t v) {
  return v == RMQ_VARIANT_LANE || v == RMQ_VARIANT_LANE8 || v == RMQ_VARIANT_LANE4;
}

// The plain-ABI build parameters that produce the hierarchy a variant reads.
inline void variant_build(uint32_t variant, int* alg, rmq_build_options* o) {
  if (variant == RMQ_VARIANT_LANE8 || variant == RMQ_VARIANT_LANE4) {
    *alg = RMQ_ALG_VECTOR_LOAD;  // R = 4 * 2^1 or 4 * 2^0
    *o = rmq_build_options{0, variant == RMQ_VARIANT_LANE8 ? 1u : 0u, 0, RMQ_BUILD_DEFAULT_CG};
  } else {
    *alg = RMQ_ALG_XXX;          // R = 2^(3 + 2); the multi-array kernels read no block scans
    *o = rmq_build_options{3, 2, variant == RMQ_VARIANT_LANE ? 0u : variant, RMQ_BUILD_NO_BLOCK_SCANS};
  }
}
constexpr uint32_t kQueues = 8;          // XCDs of an MI355X
constexpr uint32_t kTileQueries = kBlock;
constexpr uint32_t kWaveQueries = 64;
constexpr uint32_t kUnitsPerTile = kTileQueries / kWaveQueries;
#ifndef RMQ_CLAIM_UNITS
#define RMQ_CLAIM_UNITS 4
#endif
constexpr uint32_t kClaimUnits = RMQ_CLAIM_UNITS;  // units one counter increment hands to a wavefront
constexpr uint32_t kCounterStride = 64;  // uint32 words between queue counters (256 bytes)
constexpr uint32_t kCounterWords = kQueues * kCounterStride;
// Tiles one launch may hold: keeps every counter value (units plus the failed claims of all
// wavefronts) far below 2^32.
constexpr uint32_t kMaxTiles = 1u << 28;

struct QueryTile {
  uint32_t array;
  uint32_t count;  // queries in the tile (<= 256)
  uint64_t q0;     // first query index (in the query / output buffers of the launch)
};

// Zeroes the queue counters (a kernel, not a memset, so that it can sit in a captured graph).
template <uint32_t N>
__global__ void k_reset_counters(uint32_t* __restrict__ c) {
  if (threadIdx.x < N) c[threadIdx.x * kCounterStride] = 0;
}

// ---- query: persistent, tiles pulled from per-XCD queues ----------------------------------------
// One 64-query unit: queries q0 .. q0 + count - 1 (count <= 64), one per lane of the wavefront.
template <int G, bool IDX, class LT>
__device__ __forceinline__ void unit_groups(const HierView& H, const LT& T, const uint2* __restrict__ Q,
                                            uint64_t q0, uint32_t count, float* __restrict__ ov,
                                            uint32_t* __restrict__ oi) {
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t gl = lane & (G - 1);
  const uint32_t local = lane;
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
        const GroupAnswer a = solve_cl_pruned<G, kMultiRLog, true, false>(lk, ek, gl, H, T);
        v = a.v;
        i = a.idx;
      } else {
        v = value_exact<G, kMultiRLog, true, false>(lk, ek, gl, H, T, *H.zeros);
      }
      if (gl == static_cast<uint32_t>(k)) {
        my_v = v;
        my_i = i;
      }
    }
  }
  if (valid) {
    if (dev::nan_at_l(*H.zeros, H.A, l, my_v)) my_i = l;
    if (ov) ov[q0 + local] = my_v;
    if (IDX && oi) oi[q0 + local] = my_i;
  }
}

// TOP: long final ranges on the top level come from the array's sparse table when it has one
// (the lanes-only kernel, which has the registers for it).
template <int RL, bool IDX, bool TOP, class LT>
__device__ __forceinline__ void unit_lanes(const HierView& H, const LT& T, const uint2* __restrict__ Q,
                                           uint64_t q0, uint32_t count, float* __restrict__ ov,
                                           uint32_t* __restrict__ oi) {
  const uint32_t local = threadIdx.x & 63;
  if (local >= count) return;
  const uint2 p = dev::load_query(Q, q0 + local);
  if (emit_nan_at_l<IDX>(*H.zeros, H.A, p.x, q0 + local, ov, oi)) return;
  GroupAnswer a = solve_vl_pruned<RL, IDX, TOP, LT, false>(p.x, p.y + 1, H, T);
  if constexpr (!IDX) {
    if (a.v == 0.0f && !dev::zero_known(*H.zeros, a.v)) a = solve_vl_pruned<RL, true, false>(p.x, p.y + 1, H, T);
  }
  if (ov) ov[q0 + local] = a.v;
  if (IDX && oi) oi[q0 + local] = a.idx;
}

// Answers unit u (64 queries) of queue q with the variant of the unit's array. The wavefront
// first copies that array's level table into its own LDS slot `meta` (its LDS operations complete
// in order, so the previous unit's reads of the slot are done before these writes land).
template <bool IDX, bool LANES>
__device__ __forceinline__ void run_unit(uint32_t* meta, const HierView* __restrict__ views,
                                         const uint32_t* __restrict__ variant,
                                         const QueryTile* __restrict__ tiles, uint32_t first, uint32_t u,
                                         const uint2* __restrict__ Q, float* __restrict__ ov,
                                         uint32_t* __restrict__ oi) {
  const QueryTile T = tiles[first + u / kUnitsPerTile];
  const uint32_t part = (u % kUnitsPerTile) * kWaveQueries;
  if (part >= T.count) return;  // the unused quarters of a segment's last tile
  const uint32_t cnt = min(T.count - part, kWaveQueries);
  const uint64_t q0 = T.q0 + part;
  const uint32_t a = __builtin_amdgcn_readfirstlane(T.array);
  const HierView& H = views[a];
  const uint32_t lane = threadIdx.x & 63u;
  if (lane < kMetaWords) meta[lane] = lds_meta(H, lane);
  const LevelsLds L{meta};
  if constexpr (LANES) {  // every array of the launch runs a lane variant
    const uint32_t v = variant[a];
    if (v == RMQ_VARIANT_LANE4) unit_lanes<kLane4RLog, IDX, true>(H, L, Q, q0, cnt, ov, oi);
    else if (v == RMQ_VARIANT_LANE8) unit_lanes<kLane8RLog, IDX, true>(H, L, Q, q0, cnt, ov, oi);
    else unit_lanes<kMultiRLog, IDX, true>(H, L, Q, q0, cnt, ov, oi);
    return;
  }
  switch (variant[a]) {
    case RMQ_VARIANT_GROUP4: unit_groups<4, IDX>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_GROUP16: unit_groups<16, IDX>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_GROUP32: unit_groups<32, IDX>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_WAVE: unit_groups<64, IDX>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_LANE: unit_lanes<kMultiRLog, IDX, false>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_LANE8: unit_lanes<kLane8RLog, IDX, false>(H, L, Q, q0, cnt, ov, oi); break;
    case RMQ_VARIANT_LANE4: unit_lanes<kLane4RLog, IDX, false>(H, L, Q, q0, cnt, ov, oi); break;
    default: unit_groups<8, IDX>(H, L, Q, q0, cnt, ov, oi); break;
  }
}

// Wavefronts whose home is queue q (those of workgroups q, q + 8, ...) and the static share of
// each: the first half of a queue's units is split evenly among its home wavefronts, which take
// their share without touching the counter; the rest is handed out kClaimUnits at a time.
__device__ __forceinline__ void queue_shape(const uint32_t* __restrict__ qstart, uint32_t q,
                                            uint32_t& first, uint32_t& units, uint32_t& dyn0,
                                            uint32_t& share) {
  constexpr uint32_t kWaves = kBlock / kWaveQueries;
  first = qstart[q];
  units = (qstart[q + 1] - first) * kUnitsPerTile;
  const uint32_t blocks = gridDim.x > q ? (gridDim.x - q + kQueues - 1) / kQueues : 0;
  const uint32_t waves = blocks * kWaves;
  share = waves ? units / (2 * waves) : 0;
  dyn0 = share * waves;
}

// Eight wavefronts per SIMD (64 VGPRs): with every variant's solver (LANES = false) the
// occupancy is worth more than the few registers the compiler then spills (measured); LANES =
// true, for launches whose arrays all run lane variants, leaves the group solvers out and, with
// the sparse table over each array's top level (the batched build makes one for top levels of
// 64..8192 entries), fits 64 registers with 16 bytes per lane spilled in index mode; held to 7
// workgroups per CU it needs 67 and spills nothing, but ran 4% slower (4096-array workload). The kernel with every variant
// scans the top level instead: the table path would spill there.
template <bool IDX, bool LANES>
__global__ __launch_bounds__(kBlock, 8) void k_query_batch(
    const HierView* __restrict__ views, const uint32_t* __restrict__ variant,
    const QueryTile* __restrict__ tiles, const uint32_t* __restrict__ qstart,
    uint32_t* __restrict__ counter, const uint2* __restrict__ Q, float* __restrict__ ov,
    uint32_t* __restrict__ oi) {
  // Queues a wavefront of this workgroup found drained; the others skip them.
  __shared__ uint32_t s_drained;
  __shared__ uint32_t s_meta[kBlock / kWaveQueries][kMetaWords];  // each wavefront's level table
  if (threadIdx.x == 0) s_drained = 0;
  __syncthreads();
  uint32_t* meta = s_meta[threadIdx.x / kWaveQueries];
  const uint32_t lane = threadIdx.x & 63;
  const uint32_t home = blockIdx.x % kQueues;
  {
    uint32_t first, units, dyn0, share;
    queue_shape(qstart, home, first, units, dyn0, share);
    const uint32_t w = (blockIdx.x / kQueues) * (kBlock / kWaveQueries) + threadIdx.x / kWaveQueries;
    for (uint32_t u = w * share; u < (w + 1) * share; ++u)
      run_unit<IDX, LANES>(meta, views, variant, tiles, first, u, Q, ov, oi);
  }
  uint32_t queue = home;
  uint32_t tried = 0;
  while (tried < kQueues) {
    uint32_t first, units, dyn0, share;
    queue_shape(qstart, queue, first, units, dyn0, share);
    uint32_t u0 = units;
    if (lane == 0 && !(__hip_atomic_load(&s_drained, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) >> queue & 1u))
      u0 = dyn0 + atomicAdd(&counter[queue * kCounterStride], kClaimUnits);
    u0 = __builtin_amdgcn_readfirstlane(u0);
    if (u0 >= units) {  // this queue is drained: move to the next one
      if (lane == 0)
        __hip_atomic_fetch_or(&s_drained, 1u << queue, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
      queue = (queue + 1) % kQueues;
      ++tried;
      continue;
    }
    const uint32_t u1 = min(u0 + kClaimUnits, units);
    for (uint32_t u = u0; u < u1; ++u) run_unit<IDX, LANES>(meta, views, variant, tiles, first, u, Q, ov, oi);
  }
}

// Tiles of 256 queries per segment: segment a holds cnt[a] queries starting at query index
// start[a] of the launch's buffers, on array ids[a] (ids == nullptr: array a). Segments are taken
// in order of array (segments of one array adjacent, otherwise in the given order), and the
// tiles are cut into kQueues contiguous queues of about equal query count. Returns the tiles and
// the queue starts, or false when they would exceed kMaxTiles.
inline bool make_segment_tiles(uint32_t count, const uint64_t* start, const uint64_t* cnt, std::vector<QueryTile>& tiles,
                               std::vector<uint32_t>& qstart, const uint32_t* ids = nullptr) {
  uint64_t need = 0, total = 0;
  for (uint32_t a = 0; a < count; ++a) {
    need += (cnt[a] + kTileQueries - 1) / kTileQueries;
    total += cnt[a];
  }
  if (need > kMaxTiles) return false;
  std::vector<uint32_t> order(count);
  std::iota(order.begin(), order.end(), 0u);
  if (ids)
    std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) { return ids[x] < ids[y]; });
  tiles.clear();
  tiles.reserve(need);
  qstart.assign(kQueues + 1, 0);
  uint32_t queue = 0;
  uint64_t done = 0;
  for (uint32_t a : order) {
    for (uint64_t k = 0; k < cnt[a]; k += kTileQueries) {
      // queue k holds the tiles that start before k+1 eighths of all queries
      while (queue + 1 < kQueues && done * kQueues >= total * (queue + 1)) qstart[++queue] = static_cast<uint32_t>(tiles.size());
      const uint64_t c = std::min<uint64_t>(kTileQueries, cnt[a] - k);
      tiles.push_back(QueryTile{ids ? ids[a] : a, static_cast<uint32_t>(c), start[a] + k});
      done += c;
    }
  }
  while (queue + 1 < kQueues) qstart[++queue] = static_cast<uint32_t>(tiles.size());
  qstart[kQueues] = static_cast<uint32_t>(tiles.size());
  return true;
}

// The same for segments laid out back to back: segment a is [qoffsets[a], qoffsets[a + 1]).
inline bool make_query_tiles(uint32_t count, const uint64_t* qoffsets, std::vector<QueryTile>& tiles,
                             std::vector<uint32_t>& qstart, const uint32_t* ids = nullptr) {
  std::vector<uint64_t> cnt(count);
  for (uint32_t a = 0; a < count; ++a) cnt[a] = qoffsets[a + 1] - qoffsets[a];
  return make_segment_tiles(count, qoffsets, cnt.data(), tiles, qstart, ids);
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

// Counter reset + persistent query launch over the ntiles tiles of one plan. The grid is the
// resident maximum, or fewer workgroups when a small launch would leave most wavefronts with
// nothing to claim (each workgroup's first round claims kClaimUnits units per wavefront).
template <bool IDX, bool LANES>
inline void launch_batch_kernel(uint32_t small, const HierView* views, const uint32_t* variant,
                                const QueryTile* tiles, const uint32_t* qstart, uint32_t* counter,
                                const uint2* Q, float* ov, uint32_t* oi, hipStream_t s) {
  static const uint32_t grid = persistent_grid(reinterpret_cast<const void*>(k_query_batch<IDX, LANES>));
  k_query_batch<IDX, LANES><<<std::min(grid, small), kBlock, 0, s>>>(views, variant, tiles, qstart, counter, Q, ov, oi);
}

// lanes_only: every array of the launch runs a lane variant (variant_is_lane), so the
// kernel without the group solvers (fewer registers) serves it.
inline void launch_multi_query(const HierView* views, const uint32_t* variant, const QueryTile* tiles,
                               const uint32_t* qstart, uint32_t* counter, uint32_t ntiles,
                               const uint2* Q, float* ov, uint32_t* oi, hipStream_t s,
                               bool lanes_only = false) {
  constexpr uint32_t kTilesPerRound = kBlock / kWaveQueries * kClaimUnits / kUnitsPerTile;
  const uint32_t want = (ntiles + kTilesPerRound - 1) / kTilesPerRound;
  const uint32_t small = (want + kQueues - 1) / kQueues * kQueues;
  k_reset_counters<kQueues><<<1, 64, 0, s>>>(counter);
  if (oi) {
    if (lanes_only) launch_batch_kernel<true, true>(small, views, variant, tiles, qstart, counter, Q, ov, oi, s);
    else launch_batch_kernel<true, false>(small, views, variant, tiles, qstart, counter, Q, ov, oi, s);
  } else {
    if (lanes_only) launch_batch_kernel<false, true>(small, views, variant, tiles, qstart, counter, Q, ov, oi, s);
    else launch_batch_kernel<false, false>(small, views, variant, tiles, qstart, counter, Q, ov, oi, s);
  }
}

}  // namespace rmq
