// rmq_batch.hip — batched form (include/rmq_batch.h): many arrays, one build launch per level,
// one persistent query launch fed by device-side work queues, HIP-graph capture.
//
// Device layout: every array gets a HierView (the same level metadata the plain kernels read,
// rmq_device.hpp) in a table in global memory; its level buffers are slices of two arenas
// (values, leftmost indices). The query side is rmq_multi.hpp.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <new>
#include <numeric>
#include <vector>

#include "../../include/rmq_batch.h"
#include "rmq_internal.hpp"
#include "rmq_multi.hpp"

using rmq::HierView;
using rmq::kBlock;

namespace {

using rmq::QueryTile;
using rmq::kQueues;

}  // namespace

struct rmq_batch {
  uint32_t count;
  int device;
  HierView* d_views;
  uint32_t* d_variant;
  float* lv;
  uint8_t* lr;      // offset entries of every array's levels 1..
  uint32_t* zeros;  // per array: kZeroPos | kZeroNeg | kHasNaN of its level 0
  float4* tails;    // per array: the copy of level 0's last 16-byte vector, +inf padded
  uint2* tt = nullptr;  // sparse tables over the arrays' top levels (those of 64..kTopSmall entries)
  uint64_t tt_entries = 0;
  uint64_t level_floats;
  uint64_t rel_bytes;
  std::vector<uint32_t> variant;  // host copy
  std::vector<HierView> views;    // host copy of d_views
  bool lanes_only = false;        // every array runs a lane variant
};

struct rmq_batch_plan {
  uint32_t ntiles;
  QueryTile* d_tiles;
  uint32_t* d_qstart;   // [kQueues + 1] tile index where each queue starts
  uint32_t* d_counter;  // [kCounterWords] next unit of each queue, one per 256-byte line
};

struct rmq_graph {
  hipGraph_t graph;
  hipGraphExec_t exec;
};

namespace rmq {

// ---- synthetic data --------------------------------------------------------------------------
__device__ __forceinline__ float splitmix_uniform(uint64_t seed, uint64_t i) {
  uint64_t z = seed + i * 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  return static_cast<float>(static_cast<uint32_t>(z >> 40)) * (1.0f / 16777216.0f);
}

__global__ __launch_bounds__(kBlock) void k_fill_uniform(float* __restrict__ out, uint64_t n,
                                                         uint64_t seed, uint64_t start) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * kBlock;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n; i += stride) {
    out[i] = splitmix_uniform(seed, start + i);
  }
}

// ---- build: one launch per (pass, R) over every array with that R ------------------------------
// Workgroup b builds tile b - first[a] of array list[a], a the last with first[a] <= b
// (first: each listed array's first tile, ascending; first[count] = the launch's tiles).
template <int RL>
__global__ __launch_bounds__(kBlock) void k_build_pass_batch(const HierView* __restrict__ views,
                                                             const uint32_t* __restrict__ list,
                                                             const uint32_t* __restrict__ first,
                                                             uint32_t count, uint32_t k) {
  __shared__ float s_v[PassShape<RL>::LDS];
  __shared__ uint32_t s_i[PassShape<RL>::LDS];
  const uint32_t b = blockIdx.x;
  uint32_t lo = 0, hi = count;  // first[lo] <= b < first[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (first[mid] <= b) lo = mid;
    else hi = mid;
  }
  const HierView& H = views[list[lo]];
  const uint32_t m = min(PassShape<RL>::M, H.num_levels - 1 - k);
  build_pass_tile<RL>(H, k, m, b - first[lo], s_v, s_i);
}

// The sparse table over the top level of array list[blockIdx.x] (64 <= m <= kTopSmall entries),
// built in LDS by one workgroup, as the plain build's k_top_table_small does.
__global__ __launch_bounds__(kTopThreads) void k_top_table_batch(const HierView* __restrict__ views,
                                                                 const uint32_t* __restrict__ variant,
                                                                 const uint32_t* __restrict__ list) {
  __shared__ uint2 cur[kTopSmall];
  const uint32_t a = list[blockIdx.x];
  const HierView& H = views[a];
  const uint32_t top = H.num_levels - 1;
  const uint32_t m = H.size[top];
  uint32_t rows = 0;
  while ((2u << rows) <= m) ++rows;  // floor(log2 m)
  top_table_block(top == 0 ? H.A : H.lv + H.off[top], m, const_cast<uint2*>(H.tt), H.tstride, rows, cur,
                  top == 0 ? nullptr : H.lr, H.roff[top], variant_rlog(variant[a]) * top);
}

// Every array's level-0 word (cleared for the level-0 pass, or every bit when there is none)
// and its copy of level 0's last vector.
__global__ void k_init_zeros(const HierView* __restrict__ views, uint32_t count) {
  const uint32_t a = blockIdx.x * blockDim.x + threadIdx.x;
  if (a < count) {
    const HierView& H = views[a];
    *const_cast<uint32_t*>(H.zeros) = H.num_levels > 1 ? 0u : kLevel0Unknown;
    dev::fill_tail(H.A, H.size[0], const_cast<float*>(H.tail));
  }
}

// ---- the batched query without a plan (rmq_batch_query_noalloc) ----------------------------------
// Its work queue lives in the file-scope array `scratch` below instead of a plan's allocation:
// the queue counters, the queue starts, every array's first tile and the 256-query tiles, built
// on the device from the caller's query offsets by the two kernels that follow.
constexpr uint32_t kStaticArrays = RMQ_BATCH_NOALLOC_MAX_ARRAYS;
constexpr uint32_t kStaticTiles = RMQ_BATCH_NOALLOC_MAX_TILES;
constexpr size_t kScrCounters = 0;                                  // kCounterWords words
constexpr size_t kScrQstart = kScrCounters + kCounterWords * 4;     // kQueues + 1 words
constexpr size_t kScrTileStart = kScrQstart + 256;                  // kStaticArrays + 1 words
constexpr size_t kScrTiles = (kScrTileStart + 4 * (kStaticArrays + 1) + 255) & ~size_t(255);
constexpr size_t kScratchBytes = kScrTiles + kStaticTiles * sizeof(QueryTile);
constexpr uint32_t kTileThreads = 1024;

// One workgroup: tile counts per array (ceil(q_a / 256), clamped so that a wrong offset cannot
// overflow), their exclusive prefix sum (the first tile of each array, and the total at
// [count]), the queue starts (make_query_tiles's rule: queue k starts at the first tile with
// at least k/8 of all queries before it) and zeroed queue counters. A layout of more than
// kStaticTiles tiles (offsets that do not match the caller's total) becomes an empty one.
__global__ __launch_bounds__(kTileThreads) void k_static_tile_starts(const uint64_t* __restrict__ qoff,
                                                                     uint32_t count, uint8_t* __restrict__ scr) {
  uint32_t* counter = reinterpret_cast<uint32_t*>(scr + kScrCounters);
  uint32_t* qstart = reinterpret_cast<uint32_t*>(scr + kScrQstart);
  uint32_t* ts = reinterpret_cast<uint32_t*>(scr + kScrTileStart);
  __shared__ uint64_t part[kTileThreads];
  const uint32_t tid = threadIdx.x;
  const uint32_t per = (count + kTileThreads - 1) / kTileThreads;
  const uint32_t a0 = min(count, tid * per), a1 = min(count, a0 + per);
  auto tiles_of = [&](uint32_t a) -> uint64_t {
    const uint64_t c = qoff[a + 1] - qoff[a];
    const uint64_t t = c / kTileQueries + (c % kTileQueries != 0);
    return t < kStaticTiles + 1ull ? t : kStaticTiles + 1ull;
  };
  uint64_t sum = 0;
  for (uint32_t a = a0; a < a1; ++a) sum += tiles_of(a);
  part[tid] = sum;
  __syncthreads();
  for (uint32_t d = 1; d < kTileThreads; d <<= 1) {  // inclusive scan
    const uint64_t x = tid >= d ? part[tid - d] : 0ull;
    __syncthreads();
    part[tid] += x;
    __syncthreads();
  }
  const uint64_t total_tiles = part[kTileThreads - 1];
  const bool ok = total_tiles <= kStaticTiles;
  uint64_t base = part[tid] - sum;
  for (uint32_t a = a0; a < a1; ++a) {
    ts[a] = static_cast<uint32_t>(base < 0xFFFFFFFFull ? base : 0xFFFFFFFFull);
    base += tiles_of(a);
  }
  const uint32_t T = ok ? static_cast<uint32_t>(total_tiles) : 0u;
  if (tid == 0) ts[count] = T;
  if (tid < kQueues) counter[tid * kCounterStride] = 0u;
  __syncthreads();  // ts[] complete (workgroup-scope fence for global memory)
  if (tid <= kQueues) {
    uint32_t qs = T;
    const uint64_t q0 = qoff[0], total = qoff[count] - q0;
    if (tid == 0 || !ok) {
      qs = 0u;
    } else if (tid < kQueues) {
      const uint64_t x = q0 + (total * tid + kQueues - 1) / kQueues;  // the first query of queue tid
      if (x < qoff[count]) {
        uint32_t lo = 0, hi = count - 1;  // the array holding query x: smallest a with qoff[a + 1] > x
        while (lo < hi) {
          const uint32_t mid = (lo + hi) / 2;
          if (qoff[mid + 1] > x) hi = mid;
          else lo = mid + 1;
        }
        const uint64_t into = x - qoff[lo];
        qs = ts[lo] + static_cast<uint32_t>(into / kTileQueries + (into % kTileQueries != 0));
      }
    }
    qstart[tid] = qs;
  }
}

// One thread per tile t < ts[count]: its array (the last one whose first tile is <= t) and its
// queries.
__global__ __launch_bounds__(kBlock) void k_static_tiles(const uint64_t* __restrict__ qoff, uint32_t count,
                                                         uint8_t* __restrict__ scr) {
  const uint32_t* ts = reinterpret_cast<const uint32_t*>(scr + kScrTileStart);
  QueryTile* tiles = reinterpret_cast<QueryTile*>(scr + kScrTiles);
  const uint32_t t = blockIdx.x * kBlock + threadIdx.x;
  if (t >= ts[count]) return;
  uint32_t lo = 0, hi = count;  // ts[lo] <= t < ts[hi]
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) / 2;
    if (ts[mid] <= t) lo = mid;
    else hi = mid;
  }
  const uint64_t q0 = qoff[lo] + static_cast<uint64_t>(kTileQueries) * (t - ts[lo]);
  const uint64_t left = qoff[lo + 1] - q0;
  tiles[t] = QueryTile{lo, static_cast<uint32_t>(left < kTileQueries ? left : kTileQueries), q0};
}

}  // namespace rmq

// The work queue of rmq_batch_query_noalloc on each GPU: one copy per device, held until the
// process ends. Host code never uses this name as an address (it is a host-side stand-in); the
// device address comes from hipGetSymbolAddress, once per device (static_scratch below).
static __device__ __align__(256) uint8_t scratch[rmq::kScratchBytes];

namespace {

rmq_status err() { return hipGetLastError() == hipSuccess ? RMQ_OK : RMQ_EDEVICE; }

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch[kMaxDevices] = {};

// The device address of `scratch` on the current device, looked up on the first call there.
uint8_t* static_scratch() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[dev]) {
    void* p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
      (void)hipGetLastError();
      return nullptr;
    }
    g_scratch[dev] = static_cast<uint8_t*>(p);
  }
  return g_scratch[dev];
}

template <typename T>
rmq_status dmalloc(T** p, size_t count) {
  if (hipMalloc(reinterpret_cast<void**>(p), std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return RMQ_ENOMEM;
  }
  return RMQ_OK;
}

}  // namespace

extern "C" {

rmq_status rmq_fill_uniform(float* d_out, uint64_t n, uint64_t seed, uint64_t start, void* stream) {
  if (!d_out) return RMQ_EINVAL;
  if (n == 0) return RMQ_OK;
  uint64_t blocks = (n + kBlock - 1) / kBlock;
  if (blocks > 65536) blocks = 65536;
  rmq::k_fill_uniform<<<static_cast<uint32_t>(blocks), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      d_out, n, seed, start);
  return err();
}

rmq_status rmq_batch_build(const float* d_arrays, const rmq_batch_array* arrays, uint32_t count,
                           void* stream, rmq_batch** out) {
  if (!out) return RMQ_EINVAL;
  *out = nullptr;
  if (!d_arrays || !arrays || count == 0) return RMQ_EINVAL;
  if ((reinterpret_cast<uintptr_t>(d_arrays) & 15u) != 0) return RMQ_EINVAL;
  const hipStream_t s = static_cast<hipStream_t>(stream);

  // per-array plans (src/xxx.h:225-243 at R = 32) and arena offsets
  std::vector<HierView> views(count);
  std::vector<uint64_t> lv_off(count), lr_off(count);
  uint64_t arena = 0, rel_arena = 0;
  uint32_t max_levels = 1;
  for (uint32_t a = 0; a < count; ++a) {
    const rmq_batch_array& d = arrays[a];
    if (d.n == 0 || d.n > 0x7FFFFFFFu || (d.offset & 3u) != 0) return RMQ_EINVAL;
    if (!rmq::variant_ok(d.variant)) return RMQ_EUNSUPPORTED;
    uint32_t nl = 0;
    uint64_t sizes[RMQ_MAX_LEVELS], offs[RMQ_MAX_LEVELS], floats = 0;
    int alg = 0;
    rmq_build_options o;
    rmq::variant_build(d.variant, &alg, &o);
    const rmq_status st = rmq_plan(d.n, d.t, alg, &o, &nl, sizes, offs, &floats, nullptr);
    if (st != RMQ_OK) return st;
    HierView& H = views[a];
    std::memset(&H, 0, sizeof(H));
    H.num_levels = nl;
    for (uint32_t k = 0; k < nl; ++k) {
      H.size[k] = static_cast<uint32_t>(sizes[k]);
      H.off[k] = k == 0 ? 0u : static_cast<uint32_t>(offs[k]);
    }
    lv_off[a] = arena;
    arena += (floats + 31) / 32 * 32;
    lr_off[a] = rel_arena;
    rel_arena += rmq_internal_rel_layout(rmq::variant_rlog(d.variant), nl, sizes, H.roff);
    max_levels = std::max(max_levels, nl);
  }

  // sparse tables over top levels of 64..kTopSmall entries: floor(log2 m) rows of a stride of m
  // rounded up to 64 (the lanes-only query kernel reads them; the others ignore them)
  std::vector<uint64_t> tt_off(count, ~0ull);
  std::vector<uint32_t> tt_list;
  uint64_t tt_entries = 0;
  for (uint32_t a = 0; a < count; ++a) {
    const uint32_t m = views[a].size[views[a].num_levels - 1];
    if (m < 64 || m > rmq::kTopSmall) continue;
    uint32_t rows = 0;
    while ((2ull << rows) <= m) ++rows;
    views[a].tstride = (m + 63) & ~63u;
    tt_off[a] = tt_entries;
    tt_entries += static_cast<uint64_t>(rows) * views[a].tstride;
    tt_list.push_back(a);
  }

  rmq_batch* b = new (std::nothrow) rmq_batch();
  if (!b) return RMQ_ENOMEM;
  b->count = count;
  b->tt_entries = tt_entries;
  b->level_floats = arena;
  b->rel_bytes = rel_arena;
  (void)hipGetDevice(&b->device);
  rmq_status st = dmalloc(&b->lv, arena);
  if (st == RMQ_OK) st = dmalloc(&b->lr, rel_arena);
  if (st == RMQ_OK) st = dmalloc(&b->zeros, count);
  if (st == RMQ_OK) st = dmalloc(&b->tails, count);
  if (st == RMQ_OK) st = dmalloc(&b->d_views, count);
  if (st == RMQ_OK) st = dmalloc(&b->d_variant, count);
  if (st == RMQ_OK && tt_entries) st = dmalloc(&b->tt, tt_entries);
  if (st != RMQ_OK) {
    rmq_batch_free(b);
    return st;
  }
  b->variant.resize(count);
  for (uint32_t a = 0; a < count; ++a) {
    HierView& H = views[a];
    H.A = d_arrays + arrays[a].offset;
    H.lv = b->lv + lv_off[a];
    H.lr = b->lr + lr_off[a];
    H.zeros = b->zeros + a;
    H.tail = reinterpret_cast<const float*>(b->tails + a);
    H.tt = tt_off[a] == ~0ull ? nullptr : b->tt + tt_off[a];
    if (!H.tt) H.tstride = 0;
    b->variant[a] = arrays[a].variant;
  }
  b->views = views;
  b->lanes_only = std::all_of(b->variant.begin(), b->variant.end(), [](uint32_t v) { return rmq::variant_is_lane(v); });

  // build passes, grouped by reduction factor and pass (a pass from level k makes levels
  // k+1 .. k+M of every array of that R with levels above k); one device word array holds the
  // top-table list, then each group's array list and first tiles
  struct PassGroup {
    uint32_t rl, k, cnt, list, first, tiles;
  };
  constexpr uint32_t kRLogs[3] = {rmq::kMultiRLog, rmq::kLane8RLog, rmq::kLane4RLog};
  std::vector<uint32_t> words(tt_list);
  std::vector<PassGroup> groups;
  for (uint32_t rl : kRLogs) {
    const uint32_t M = rmq::kPassTileLog / rl;  // PassShape<rl>::M
    for (uint32_t k = 0; k + 1 < max_levels; k += M) {
      std::vector<uint32_t> list, first;
      uint64_t tiles = 0;
      for (uint32_t a = 0; a < count; ++a) {
        if (rmq::variant_rlog(arrays[a].variant) != rl || views[a].num_levels <= k + 1) continue;
        list.push_back(a);
        first.push_back(static_cast<uint32_t>(tiles));
        tiles += (static_cast<uint64_t>(views[a].size[k]) + (1u << rmq::kPassTileLog) - 1) >> rmq::kPassTileLog;
      }
      if (list.empty()) continue;
      if (tiles > 0xFFFFFFFFull / kBlock) {  // one launch's threads must fit in 32 bits
        rmq_batch_free(b);
        return RMQ_EINVAL;
      }
      first.push_back(static_cast<uint32_t>(tiles));
      const PassGroup g{rl, k, static_cast<uint32_t>(list.size()), static_cast<uint32_t>(words.size()),
                        static_cast<uint32_t>(words.size() + list.size()), static_cast<uint32_t>(tiles)};
      words.insert(words.end(), list.begin(), list.end());
      words.insert(words.end(), first.begin(), first.end());
      groups.push_back(g);
    }
  }
  uint32_t* d_words = nullptr;
  if (!words.empty()) {
    st = dmalloc(&d_words, words.size());
    if (st != RMQ_OK) {
      rmq_batch_free(b);
      return st;
    }
  }
  bool ok = hipMemcpyAsync(b->d_views, views.data(), count * sizeof(HierView), hipMemcpyHostToDevice, s) == hipSuccess &&
            hipMemcpyAsync(b->d_variant, b->variant.data(), count * sizeof(uint32_t), hipMemcpyHostToDevice, s) == hipSuccess;
  if (ok && !words.empty())
    ok = hipMemcpyAsync(d_words, words.data(), words.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess;
  if (ok) {
    rmq::k_init_zeros<<<(count + 255) / 256, 256, 0, s>>>(b->d_views, count);
    ok = hipGetLastError() == hipSuccess;
  }
  // passes in order of k within each R (a pass reads the level the one before it made)
  for (const PassGroup& g : groups) {
    if (!ok) break;
    const uint32_t* list = d_words + g.list;
    const uint32_t* first = d_words + g.first;
    switch (g.rl) {
      case rmq::kMultiRLog:
        rmq::k_build_pass_batch<rmq::kMultiRLog><<<g.tiles, kBlock, 0, s>>>(b->d_views, list, first, g.cnt, g.k);
        break;
      case rmq::kLane8RLog:
        rmq::k_build_pass_batch<rmq::kLane8RLog><<<g.tiles, kBlock, 0, s>>>(b->d_views, list, first, g.cnt, g.k);
        break;
      default:
        rmq::k_build_pass_batch<rmq::kLane4RLog><<<g.tiles, kBlock, 0, s>>>(b->d_views, list, first, g.cnt, g.k);
        break;
    }
    ok = hipGetLastError() == hipSuccess;
  }
  const uint32_t* d_tt_list = d_words;
  if (ok && !tt_list.empty()) {  // after every level: the tables read the top levels
    rmq::k_top_table_batch<<<static_cast<uint32_t>(tt_list.size()), rmq::kTopThreads, 0, s>>>(b->d_views, b->d_variant,
                                                                                              d_tt_list);
    ok = hipGetLastError() == hipSuccess;
  }
  // the host vectors die here: wait for the copies (and the build) before returning
  ok = ok && hipStreamSynchronize(s) == hipSuccess;
  if (d_words) (void)hipFree(d_words);
  if (!ok) {
    rmq_batch_free(b);
    return RMQ_EDEVICE;
  }
  *out = b;
  return RMQ_OK;
}

rmq_status rmq_batch_plan_queries(const rmq_batch* b, const uint64_t* qoffsets, void* stream,
                                  rmq_batch_plan** out) {
  if (!b || !qoffsets || !out) return RMQ_EINVAL;
  *out = nullptr;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  for (uint32_t a = 0; a < b->count; ++a)
    if (qoffsets[a + 1] < qoffsets[a]) return RMQ_EINVAL;
  std::vector<QueryTile> tiles;
  std::vector<uint32_t> qstart;
  if (!rmq::make_query_tiles(b->count, qoffsets, tiles, qstart)) return RMQ_EINVAL;
  rmq_batch_plan* p = new (std::nothrow) rmq_batch_plan();
  if (!p) return RMQ_ENOMEM;
  p->ntiles = static_cast<uint32_t>(tiles.size());
  rmq_status st = dmalloc(&p->d_tiles, tiles.size());
  if (st == RMQ_OK) st = dmalloc(&p->d_qstart, kQueues + 1);
  if (st == RMQ_OK) st = dmalloc(&p->d_counter, rmq::kCounterWords);
  if (st != RMQ_OK) {
    rmq_batch_plan_free(p);
    return st;
  }
  bool ok = hipMemcpyAsync(p->d_qstart, qstart.data(), qstart.size() * 4, hipMemcpyHostToDevice, s) == hipSuccess;
  if (ok && !tiles.empty())
    ok = hipMemcpyAsync(p->d_tiles, tiles.data(), tiles.size() * sizeof(QueryTile), hipMemcpyHostToDevice, s) == hipSuccess;
  ok = ok && hipStreamSynchronize(s) == hipSuccess;
  if (!ok) {
    rmq_batch_plan_free(p);
    return RMQ_EDEVICE;
  }
  *out = p;
  return RMQ_OK;
}

rmq_status rmq_batch_query(const rmq_batch* b, const rmq_batch_plan* p, const uint32_t* d_lr,
                           float* d_val, uint32_t* d_idx, void* stream) {
  if (!b || !p) return RMQ_EINVAL;
  if (p->ntiles == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx) || (reinterpret_cast<uintptr_t>(d_lr) & 7u) != 0) return RMQ_EINVAL;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  const uint2* Q = reinterpret_cast<const uint2*>(d_lr);
  rmq::launch_multi_query(b->d_views, b->d_variant, p->d_tiles, p->d_qstart, p->d_counter, p->ntiles,
                          Q, d_val, d_idx, s, b->lanes_only);
  return err();
}

rmq_status rmq_batch_query_noalloc(const rmq_batch* b, const uint64_t* d_qoffsets, uint64_t total_queries,
                                   const uint32_t* d_lr, float* d_val, uint32_t* d_idx, void* stream) {
  if (!b || !d_qoffsets || (reinterpret_cast<uintptr_t>(d_qoffsets) & 7u) != 0) return RMQ_EINVAL;
  if (total_queries == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx) || (reinterpret_cast<uintptr_t>(d_lr) & 7u) != 0) return RMQ_EINVAL;
  // every array has at most q_a / 256 + 1 tiles
  const uint64_t max_tiles = total_queries / rmq::kTileQueries + b->count;
  if (b->count > rmq::kStaticArrays || max_tiles > rmq::kStaticTiles) return RMQ_EINVAL;
  uint8_t* scr = static_scratch();
  if (!scr) return RMQ_EDEVICE;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  rmq::k_static_tile_starts<<<1, rmq::kTileThreads, 0, s>>>(d_qoffsets, b->count, scr);
  rmq::k_static_tiles<<<static_cast<uint32_t>((max_tiles + kBlock - 1) / kBlock), kBlock, 0, s>>>(d_qoffsets, b->count,
                                                                                                  scr);
  rmq::launch_multi_query(b->d_views, b->d_variant, reinterpret_cast<const QueryTile*>(scr + rmq::kScrTiles),
                          reinterpret_cast<const uint32_t*>(scr + rmq::kScrQstart),
                          reinterpret_cast<uint32_t*>(scr + rmq::kScrCounters), static_cast<uint32_t>(max_tiles),
                          reinterpret_cast<const uint2*>(d_lr), d_val, d_idx, s, b->lanes_only);
  return err();
}

rmq_status rmq_batch_capture(const rmq_batch* b, const rmq_batch_plan* p, const uint32_t* d_lr,
                             float* d_val, uint32_t* d_idx, void* stream, rmq_graph** out) {
  if (!out) return RMQ_EINVAL;
  *out = nullptr;
  (void)stream;
  hipStream_t cs;
  if (hipStreamCreateWithFlags(&cs, hipStreamNonBlocking) != hipSuccess) return RMQ_EDEVICE;
  rmq_graph* g = new (std::nothrow) rmq_graph();
  if (!g) {
    (void)hipStreamDestroy(cs);
    return RMQ_ENOMEM;
  }
  rmq_status st = RMQ_OK;
  if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) st = RMQ_EDEVICE;
  if (st == RMQ_OK) {
    st = rmq_batch_query(b, p, d_lr, d_val, d_idx, cs);
    hipGraph_t graph = nullptr;
    if (hipStreamEndCapture(cs, &graph) != hipSuccess) st = RMQ_EDEVICE;
    g->graph = graph;
  }
  if (st == RMQ_OK && hipGraphInstantiate(&g->exec, g->graph, nullptr, nullptr, 0) != hipSuccess)
    st = RMQ_EDEVICE;
  (void)hipStreamDestroy(cs);
  if (st != RMQ_OK) {
    rmq_graph_free(g);
    return st;
  }
  *out = g;
  return RMQ_OK;
}

rmq_status rmq_graph_launch(const rmq_graph* g, void* stream) {
  if (!g || !g->exec) return RMQ_EINVAL;
  return hipGraphLaunch(g->exec, static_cast<hipStream_t>(stream)) == hipSuccess ? RMQ_OK : RMQ_EDEVICE;
}

void rmq_graph_free(rmq_graph* g) {
  if (!g) return;
  if (g->exec) (void)hipGraphExecDestroy(g->exec);
  if (g->graph) (void)hipGraphDestroy(g->graph);
  delete g;
}

void rmq_batch_plan_free(rmq_batch_plan* p) {
  if (!p) return;
  if (p->d_tiles) (void)hipFree(p->d_tiles);
  if (p->d_qstart) (void)hipFree(p->d_qstart);
  if (p->d_counter) (void)hipFree(p->d_counter);
  delete p;
}

void rmq_batch_free(rmq_batch* b) {
  if (!b) return;
  if (b->lv) (void)hipFree(b->lv);
  if (b->zeros) (void)hipFree(b->zeros);
  if (b->tails) (void)hipFree(b->tails);
  if (b->lr) (void)hipFree(b->lr);
  if (b->d_views) (void)hipFree(b->d_views);
  if (b->d_variant) (void)hipFree(b->d_variant);
  if (b->tt) (void)hipFree(b->tt);
  delete b;
}

uint64_t rmq_batch_level_bytes(const rmq_batch* b) {
  return b ? b->level_floats * 4 + b->rel_bytes + b->tt_entries * sizeof(uint2) : 0;
}

rmq_status rmq_batch_read_levels(const rmq_batch* b, uint32_t array, float* d_val, uint32_t* d_idx,
                                 void* stream) {
  if (!b || array >= b->count || (!d_val && !d_idx)) return RMQ_EINVAL;
  return rmq_internal_read_levels(b->views[array], rmq::variant_rlog(b->variant[array]), d_val, d_idx, stream);
}

}  // extern "C"
