// rmq_api.hip — the C ABI of librmq.so (include/rmq.h): plan, build, query, free, GPU BASE.
//
// Host-side restatement of the reference's per-algorithm drivers xxx_rmq (src/xxx.h:203-298),
// xxx_rmq_multi_load (src/xxx.h:530-627) and hierarchical_vector_load (src/xxx.h:932-1040),
// split into build / query / free, stream-ordered and returning status codes.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cstdint>
#include <cstring>
#include <new>

#include "../../include/rmq.h"
#include "rmq_long.hpp"
#include "rmq_internal.hpp"
#include "rmq_kernels.hpp"

using rmq::HierView;
using rmq::kBlock;

namespace rmq {
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

// The hierarchy's small words, set from the stream (a kernel, not a memset, so that a captured
// build stays valid): w[0] = the zero-sign word, w[4..8) = the copy of level 0's last vector.
__global__ void k_init_words(uint32_t* __restrict__ w, uint32_t zeros, const float* __restrict__ A, uint32_t n) {
  w[0] = zeros;
  dev::fill_tail(A, n, reinterpret_cast<float*>(w + 4));
}

// Level read-back (rmq_read_levels): entry j of level k >= 1 to out_v[off[k] + j], its leftmost
// level-0 index (j << R_log k plus the offset entry) to out_i[off[k] + j].
__global__ __launch_bounds__(kBlock) void k_read_level(HierView H, uint32_t k, uint32_t span_log,
                                                       float* __restrict__ out_v, uint32_t* __restrict__ out_i) {
  for (uint32_t j = blockIdx.x * kBlock + threadIdx.x; j < H.size[k]; j += gridDim.x * kBlock) {
    const uint32_t o = H.off[k] + j;
    if (out_v) out_v[o] = H.lv[o];
    if (out_i) out_i[o] = (span_log >= 32u ? 0u : j << span_log) + dev::rel_load(H.lr, H.roff[k], j, span_log);
  }
}

// Query validation (RMQ_BUILD_CHECK_QUERIES): *bad becomes nonzero when a query is not
// 0 <= l <= r < n. k_set_word clears it first (a kernel, so that the pair can be captured).
__global__ void k_set_word(uint32_t* __restrict__ w, uint32_t v) { *w = v; }

template <class P>
__global__ __launch_bounds__(kBlock) void k_check_queries(const P* __restrict__ Q, uint64_t q, uint64_t n,
                                                          uint32_t* __restrict__ bad) {
  const uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x;
  bool b = false;
  if (i < q) {
    const P p = Q[i];
    b = !(p.x <= p.y && static_cast<uint64_t>(p.y) < n);
  }
  if (__ballot(b) != 0 && (threadIdx.x & 63u) == 0u) atomicOr(bad, 1u);
}

// Row j >= 1 of the sparse table over the top level top[0, m): entry i (i + 2^j <= m) is the
// leftmost minimum of [i, i + 2^j), from two entries of row j - 1 (row 1 from the level itself).
// Entries past m - 2^j are never read and stay unwritten.
__global__ __launch_bounds__(kBlock) void k_top_row(const float* __restrict__ top, uint32_t m,
                                                    uint2* __restrict__ tt, uint32_t stride, uint32_t j,
                                                    const uint8_t* __restrict__ lr, uint32_t roff, uint32_t sk) {
  const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
  const uint32_t h = 1u << (j - 1u);
  if (static_cast<uint64_t>(i) + 2u * h > m) return;
  uint2 a, b;
  if (j == 1) {
    a = make_uint2(dev::f2u(top[i]), level0_index(lr, roff, i, sk));
    b = make_uint2(dev::f2u(top[i + 1]), level0_index(lr, roff, i + 1, sk));
  } else {
    const uint2* prev = tt + static_cast<size_t>(j - 2u) * stride;
    a = prev[i];
    b = prev[i + h];
  }
  tt[static_cast<size_t>(j - 1u) * stride + i] = dev::top_take_right(dev::u2f(a.x), dev::u2f(b.x)) ? b : a;
}

// The whole table of a top level of at most kTopSmall entries in one workgroup.
__global__ __launch_bounds__(kTopThreads) void k_top_table_small(const float* __restrict__ top, uint32_t m,
                                                                 uint2* __restrict__ tt, uint32_t stride,
                                                                 uint32_t rows, const uint8_t* __restrict__ lr,
                                                                 uint32_t roff, uint32_t sk) {
  __shared__ uint2 cur[kTopSmall];
  top_table_block(top, m, tt, stride, rows, cur, lr, roff, sk);
}

}  // namespace rmq

struct rmq_hierarchy {
  int alg;
  int device;
  uint32_t R_log;
  uint32_t group_lanes;
  uint32_t flags;
  uint64_t n;
  uint64_t t;
  uint64_t buffer_floats;
  uint8_t* mem = nullptr;  // the one device allocation the pointers below live in
  uint64_t mem_bytes = 0;
  float* lv;
  uint8_t* lr;      // offset entries of levels 1.. (HierView::lr)
  uint64_t rel_bytes = 0;
  uint32_t* zeros;  // device word: kZeroPos | kZeroNeg | kHasNaN of level 0
  uint2* tt = nullptr;  // sparse table over the top level (null when not built)
  HierView view;
  // 64-bit form: chunks of 2^chunk_log elements, each its own hierarchy (chunk_log = 0: one
  // hierarchy over the whole array). views / cmin / cidx are device tables: the chunks' views
  // and every chunk's minimum.
  uint32_t chunk_log = 0;
  uint32_t chunks = 0;
  uint64_t msize[RMQ_MAX_LEVELS] = {};  // the reference's metadata for the whole array
  uint64_t moff[RMQ_MAX_LEVELS] = {};
  rmq_hierarchy** part = nullptr;
  HierView* views = nullptr;
  HierView* self_view = nullptr;  // one-piece hierarchy: its own view in device memory
  float* cmin = nullptr;
  uint64_t* cidx = nullptr;
};

namespace {

constexpr uint64_t kMaxN = 0x7FFFFFFFull;  // int32 n, as the reference's CmdArgs (src/tools.h:30)
constexpr uint64_t kMaxN64 = 1ull << 38;    // the 64-bit form (nType = uint64_t, src/tools.h:27-29)
constexpr uint32_t kChunkLog = 30;          // chunk size of the 64-bit form
constexpr uint64_t kMaxGrid = 1ull << 20;
constexpr int kMaxDevices = 64;  // devices a process's per-device launch shapes are kept for

// Most bytes of block scans a hierarchy keeps (RMQ_BUILD_BLOCK_SCANS), from the top level down:
// at n = 2^30, R = 32 the scans of level 1 (512 MB, cold) were slower than its vectors, those of
// levels 2.. faster; at n = 2^26 all of them (33 MB) pay.
constexpr uint64_t kScanBytes = 64ull << 20;
// Queries one launch of a one-lane-per-query kernel answers (kMaxGrid workgroups of kBlock
// lanes); rmq_query and rmq_gpu_base split larger sets into launches of this many.
constexpr uint64_t kLaunchQueries = kMaxGrid * kBlock;
// The top level's sparse table is built when the top level holds at least kTopTableMin
// entries and the table fits in kTopTableMax entries of 8 bytes (64 MiB).
constexpr uint64_t kTopTableMin = 64;
constexpr uint64_t kTopTableMax = 1ull << 23;

struct Plan {
  uint32_t R_log = 0;
  uint32_t G = 0;
  uint32_t num_levels = 0;
  uint64_t size[RMQ_MAX_LEVELS] = {};
  uint64_t off[RMQ_MAX_LEVELS] = {};
  uint64_t buffer_floats = 0;
};

bool alg_supported(int alg) {
  return alg == RMQ_ALG_XXX || alg == RMQ_ALG_VECTOR_LOAD || alg == RMQ_ALG_XXX_MULTI_LOAD;
}

// Metadata recurrence of src/xxx.h:225-243 and the t checks of src/xxx.h:211-217 / 938-944.
rmq_status make_plan(uint64_t n, uint64_t t, int alg, const rmq_build_options* opt, Plan* P) {
  if (!alg_supported(alg)) return RMQ_EINVAL;
  if (n == 0 || n > kMaxN64 || t == 0) return RMQ_EINVAL;
  uint32_t g = 3, c = 2;  // XXX_CG_SIZE_LOG, XXX_CG_AMOUNT_LOG defaults (CMakeLists.txt:61-67)
  uint32_t G = 0;
  if (opt) {
    const bool literal = (opt->flags & RMQ_BUILD_DEFAULT_CG) != 0;
    if (opt->cg_size_log || literal) g = opt->cg_size_log;
    if (opt->cg_amount_log || literal) c = opt->cg_amount_log;
    G = opt->group_lanes;
  }
  uint32_t R_log;
  if (alg == RMQ_ALG_VECTOR_LOAD) {
    R_log = c + 2;  // src/xxx.h:934
    if (R_log < 2 || R_log > 6) return RMQ_EUNSUPPORTED;
    if (t < (1ull << R_log)) return RMQ_ETHRESHOLD;
  } else {
    R_log = g + c;  // src/xxx.h:208
    if (R_log < 3 || R_log > 7) return RMQ_EUNSUPPORTED;
    if (t < 2 * (1ull << R_log)) return RMQ_ETHRESHOLD;
    if (G == 0) {
      // a lane per query where it runs (measured fastest for uniform, lognormal and mixed
      // lengths once long top-level ranges come from the sparse table); else one 16-byte
      // vector per lane per block
      const bool full = opt && (opt->flags & RMQ_BUILD_FULL);
      const uint32_t v = (1u << R_log) / 4;
      G = (R_log <= 6 && !full) ? 1 : (v < 4 ? 4 : (v > 16 ? 16 : v));
    }
    if (G == 1) {  // a lane per query: the lane solver on this R = 2^(g+c) hierarchy
      if (R_log > 6 || (opt && (opt->flags & RMQ_BUILD_FULL))) return RMQ_EUNSUPPORTED;
    } else if (G != 4 && G != 8 && G != 16 && G != 32 && G != 64) {
      return RMQ_EUNSUPPORTED;
    }
  }
  P->R_log = R_log;
  P->G = G;
  const uint64_t R = 1ull << R_log;
  uint64_t size = n;
  P->num_levels = 1;
  P->off[0] = 0xFFFFFFFFull;  // ~uint32_t(0), src/xxx.h:229
  P->size[0] = size;
  uint64_t offset = 0;
  while (true) {
    if (P->num_levels >= RMQ_MAX_LEVELS || size <= t) break;
    size = (size + R - 1) >> R_log;
    offset = (offset + RMQ_LEVEL_ALIGNMENT - 1) & ~static_cast<uint64_t>(RMQ_LEVEL_ALIGNMENT - 1);
    P->off[P->num_levels] = offset;
    P->size[P->num_levels] = size;
    P->num_levels++;
    offset += size;
  }
  P->buffer_floats = offset;
  return RMQ_OK;
}

uint32_t grid_for(uint64_t threads) {
  uint64_t b = (threads + kBlock - 1) / kBlock;
  if (b == 0) b = 1;
  if (b > kMaxGrid) b = kMaxGrid;
  return static_cast<uint32_t>(b);
}

rmq_status last_error() { return hipGetLastError() == hipSuccess ? RMQ_OK : RMQ_EDEVICE; }

// ---- build dispatch -----------------------------------------------------------------------
// Levels 1.. of H (sizes, offsets and pointers set; level 0's word cleared), several levels per
// pass (rmq::k_build_pass): a pass from level k makes levels k+1 .. k+M.
template <int RL>
rmq_status launch_passes(const HierView& H, hipStream_t s) {
  using P = rmq::PassShape<RL>;
  for (uint32_t k = 0; k + 1 < H.num_levels; k += P::M) {
    const uint32_t m = std::min<uint32_t>(P::M, H.num_levels - 1 - k);
    const uint32_t tiles = static_cast<uint32_t>((static_cast<uint64_t>(H.size[k]) + (1u << P::SLOG) - 1) >> P::SLOG);
    rmq::k_build_pass<RL><<<tiles, kBlock, 0, s>>>(H, k, m);
    if (last_error() != RMQ_OK) return RMQ_EDEVICE;
  }
  return RMQ_OK;
}

rmq_status build_levels(uint32_t R_log, const HierView& H, hipStream_t s) {
  switch (R_log) {
    case 2: return launch_passes<2>(H, s);
    case 3: return launch_passes<3>(H, s);
    case 4: return launch_passes<4>(H, s);
    case 5: return launch_passes<5>(H, s);
    case 6: return launch_passes<6>(H, s);
    case 7: return launch_passes<7>(H, s);
    default: return RMQ_EUNSUPPORTED;
  }
}

template <int RL>
rmq_status launch_block_scans(const HierView& H, hipStream_t s) {
  for (uint32_t k = H.ps_lo; k < H.num_levels; ++k) {
    if constexpr (RL <= 6) {  // a lane per entry
      rmq::k_block_scans_lanes<RL><<<(H.size[k] + kBlock - 1) / kBlock, kBlock, 0, s>>>(H, k, const_cast<uint4*>(H.ps));
    } else {  // a thread per block
      const uint32_t blocks = (H.size[k] + (1u << RL) - 1u) >> RL;
      rmq::k_block_scans<RL><<<(blocks + kBlock - 1) / kBlock, kBlock, 0, s>>>(H, k, const_cast<uint4*>(H.ps));
    }
  }
  return last_error();
}

rmq_status build_block_scans(uint32_t R_log, const HierView& H, hipStream_t s) {
  switch (R_log) {
    case 2: return launch_block_scans<2>(H, s);
    case 3: return launch_block_scans<3>(H, s);
    case 4: return launch_block_scans<4>(H, s);
    case 5: return launch_block_scans<5>(H, s);
    case 6: return launch_block_scans<6>(H, s);
    case 7: return launch_block_scans<7>(H, s);
    default: return RMQ_EUNSUPPORTED;
  }
}

// Byte offsets of the levels' offset entries (rel_bytes per level, 16-byte aligned, four entries
// of padding each); returns the total bytes.
uint64_t rel_layout(uint32_t R_log, uint32_t num_levels, const uint64_t* size, uint32_t* roff) {
  uint64_t b = 0;
  for (uint32_t k = 0; k < rmq::kMaxLevels; ++k) roff[k] = 0;
  for (uint32_t k = 1; k < num_levels; ++k) {
    roff[k] = static_cast<uint32_t>(b);
    b += ((size[k] + 4) * rmq::rel_bytes(R_log * k) + 15) & ~15ull;
  }
  return b;
}

// ---- query dispatch -----------------------------------------------------------------------
// P = pruned evaluation order (default) or every piece read (RMQ_BUILD_FULL).
// PS: the hierarchy has block scans (read by the pruned order only).
template <int G, int RL, bool P, bool PS>
rmq_status launch_cl(int alg, const uint2* Q, uint32_t q, float* ov, uint32_t* oi,
                     const HierView& H, hipStream_t s) {
  const bool idx = oi != nullptr;
  if (alg == RMQ_ALG_XXX) {
    const uint32_t grid = grid_for(q);
    if (static_cast<uint64_t>(grid) * kBlock < q) return RMQ_EINVAL;
    if (idx) rmq::k_query_xxx<G, RL, P, true, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_xxx<G, RL, P, false, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  } else {
    const uint32_t grid = grid_for(static_cast<uint64_t>(q) * G);
    if (idx) rmq::k_query_multi<G, RL, P, true, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_multi<G, RL, P, false, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  }
  return last_error();
}

template <int G, int RL>
rmq_status launch_cl_order(bool full, int alg, const uint2* Q, uint32_t q, float* ov, uint32_t* oi,
                           const HierView& H, hipStream_t s) {
  if (full) return launch_cl<G, RL, false, false>(alg, Q, q, ov, oi, H, s);
  if (H.ps) return launch_cl<G, RL, true, true>(alg, Q, q, ov, oi, H, s);
  return launch_cl<G, RL, true, false>(alg, Q, q, ov, oi, H, s);
}

template <int RL>
rmq_status launch_cl_g(uint32_t G, bool full, int alg, const uint2* Q, uint32_t q, float* ov,
                       uint32_t* oi, const HierView& H, hipStream_t s) {
  switch (G) {
    case 4: return launch_cl_order<4, RL>(full, alg, Q, q, ov, oi, H, s);
    case 8: return launch_cl_order<8, RL>(full, alg, Q, q, ov, oi, H, s);
    case 16: return launch_cl_order<16, RL>(full, alg, Q, q, ov, oi, H, s);
    case 32: return launch_cl_order<32, RL>(full, alg, Q, q, ov, oi, H, s);
    case 64: return launch_cl_order<64, RL>(full, alg, Q, q, ov, oi, H, s);  // a wave per query
    default: return RMQ_EUNSUPPORTED;
  }
}

// The pruned lane kernels, W0: n % 4 == 0, so level-0 vector loads need no tail select; PS: the
// hierarchy has block scans.
// The LDS-top form of the paired kernel: a persistent grid of kTopBlock-lane workgroups, as many
// as fit the CUs with the top level's LDS image.
template <int RL, bool IDX, bool W0, bool PS>
void launch_vl2t(const uint2* Q, uint32_t q, float* ov, uint32_t* oi, const HierView& H, hipStream_t s) {
  const uint32_t m = H.size[H.num_levels - 1];
  const uint32_t rows = 31u - static_cast<uint32_t>(__builtin_clz(m));
  const uint32_t bytes = rmq::lds_top_bytes(m, rows);
  const void* k = reinterpret_cast<const void*>(rmq::k_query_vl2t<RL, IDX, W0, PS>);
  // per device, on its first launch there: the LDS limit raised past 64 KiB, and as many
  // workgroups as its CUs hold with the largest image (concurrent first calls compute the same)
  static std::atomic<uint32_t> grids[kMaxDevices];
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) dev = 0;
  uint32_t grid = grids[dev].load(std::memory_order_relaxed);
  if (!grid) {
    (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, 64 * 1024 * 2);
    int cus = 0, per_cu = 0;
    if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 256;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, rmq::kTopBlock, rmq::lds_top_bytes(rmq::kLdsTopMax, 10)) !=
            hipSuccess || per_cu < 1)
      per_cu = 1;
    grid = static_cast<uint32_t>(cus * per_cu);
    grids[dev].store(grid, std::memory_order_relaxed);
  }
  rmq::k_query_vl2t<RL, IDX, W0, PS><<<grid, rmq::kTopBlock, bytes, s>>>(Q, q, ov, oi, H);
}

template <int RL, bool W0, bool PS>
void launch_vl_pruned(bool packed, bool paired, uint32_t grid, const uint2* Q, uint32_t q, float* ov,
                      uint32_t* oi, const HierView& H, hipStream_t s) {
  if constexpr (RL == 2) {  // the LDS-top form is built for R = 4 only
    const uint32_t m = H.size[H.num_levels - 1];
    if (paired && !packed && H.tt && H.num_levels > 1 && m >= (1u << rmq::kLdsTopRow0) && m <= rmq::kLdsTopMax) {
      if (oi) launch_vl2t<RL, true, W0, PS>(Q, q, ov, oi, H, s);
      else launch_vl2t<RL, false, W0, PS>(Q, q, ov, oi, H, s);
      return;
    }
  }
  if (paired && packed) {
    if (oi) rmq::k_query_vl2c<RL, true, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_vl2c<RL, false, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  } else if (paired) {
    if (oi) rmq::k_query_vl2<RL, true, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_vl2<RL, false, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  } else if (packed) {
    if (oi) rmq::k_query_vlc<RL, true, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_vlc<RL, false, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  } else {
    if (oi) rmq::k_query_vl<RL, true, true, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_vl<RL, true, false, W0, PS><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  }
}

template <int RL, bool W0>
void launch_vl_ps(bool packed, bool paired, uint32_t grid, const uint2* Q, uint32_t q, float* ov,
                  uint32_t* oi, const HierView& H, hipStream_t s) {
  if (H.ps) launch_vl_pruned<RL, W0, true>(packed, paired, grid, Q, q, ov, oi, H, s);
  else launch_vl_pruned<RL, W0, false>(packed, paired, grid, Q, q, ov, oi, H, s);
}

template <int RL>
rmq_status launch_vl(bool full, bool packed, const uint2* Q, uint32_t q, float* ov, uint32_t* oi,
                     const HierView& H, hipStream_t s, bool paired = false) {
  paired = paired && !full;  // two queries per lane
  const uint64_t lanes = paired ? (static_cast<uint64_t>(q) + 1) / 2 : q;
  const uint32_t grid = grid_for(lanes);
  if (static_cast<uint64_t>(grid) * kBlock < lanes) return RMQ_EINVAL;
  if (full) {
    if (oi) rmq::k_query_vl<RL, false, true><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
    else rmq::k_query_vl<RL, false, false><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, H);
  } else if ((H.size[0] & 3u) == 0) {
    launch_vl_ps<RL, true>(packed, paired, grid, Q, q, ov, oi, H, s);
  } else {
    launch_vl_ps<RL, false>(packed, paired, grid, Q, q, ov, oi, H, s);
  }
  return last_error();
}

// RMQ_BUILD_CHECK_QUERIES: every query inside [0, n) with l <= r, checked on the device before
// any query kernel runs; the word lives in the hierarchy's words (a chunked one uses its first
// chunk's). Synchronises the stream.
template <class P>
rmq_status check_queries(const rmq_hierarchy* h, const P* Q, uint64_t q, hipStream_t s) {
  uint32_t* w = (h->chunk_log ? h->part[0]->zeros : h->zeros) + 8;
  const uint64_t grid = (q + kBlock - 1) / kBlock;
  if (grid > 0xFFFFFFFFull) return RMQ_EINVAL;
  rmq::k_set_word<<<1, 1, 0, s>>>(w, 0u);
  rmq::k_check_queries<P><<<static_cast<uint32_t>(grid), kBlock, 0, s>>>(Q, q, h->n, w);
  if (last_error() != RMQ_OK) return RMQ_EDEVICE;
  uint32_t bad = 0;
  if (hipMemcpyAsync(&bad, w, sizeof(bad), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess) {
    (void)hipGetLastError();
    return RMQ_EDEVICE;
  }
  return bad ? RMQ_EINVAL : RMQ_OK;
}

rmq_status query_dispatch(const rmq_hierarchy* h, const uint2* Q, uint32_t q, float* ov,
                          uint32_t* oi, hipStream_t s) {
  const HierView& H = h->view;
  const bool packed = (h->flags & RMQ_BUILD_PACK_DESCENTS) != 0;
  const bool paired = (h->flags & RMQ_BUILD_PAIRED) != 0;
  if (h->alg == RMQ_ALG_VECTOR_LOAD) {
    const bool full = (h->flags & RMQ_BUILD_FULL) != 0;
    switch (h->R_log) {
      case 2: return launch_vl<2>(full, packed, Q, q, ov, oi, H, s, paired);
      case 3: return launch_vl<3>(full, packed, Q, q, ov, oi, H, s, paired);
      case 4: return launch_vl<4>(full, packed, Q, q, ov, oi, H, s, paired);
      case 5: return launch_vl<5>(full, packed, Q, q, ov, oi, H, s, paired);
      case 6: return launch_vl<6>(full, packed, Q, q, ov, oi, H, s, paired);
      default: return RMQ_EUNSUPPORTED;
    }
  }
  if (h->group_lanes == 1) {  // algs 16 / 24 with a lane per query (R <= 64, pruned order)
    switch (h->R_log) {
      case 3: return launch_vl<3>(false, packed, Q, q, ov, oi, H, s, paired);
      case 4: return launch_vl<4>(false, packed, Q, q, ov, oi, H, s, paired);
      case 5: return launch_vl<5>(false, packed, Q, q, ov, oi, H, s, paired);
      case 6: return launch_vl<6>(false, packed, Q, q, ov, oi, H, s, paired);
      default: return RMQ_EUNSUPPORTED;
    }
  }
  switch (h->R_log) {
    case 3: return launch_cl_g<3>(h->group_lanes, (h->flags & RMQ_BUILD_FULL) != 0, h->alg, Q, q, ov, oi, H, s);
    case 4: return launch_cl_g<4>(h->group_lanes, (h->flags & RMQ_BUILD_FULL) != 0, h->alg, Q, q, ov, oi, H, s);
    case 5: return launch_cl_g<5>(h->group_lanes, (h->flags & RMQ_BUILD_FULL) != 0, h->alg, Q, q, ov, oi, H, s);
    case 6: return launch_cl_g<6>(h->group_lanes, (h->flags & RMQ_BUILD_FULL) != 0, h->alg, Q, q, ov, oi, H, s);
    case 7: return launch_cl_g<7>(h->group_lanes, (h->flags & RMQ_BUILD_FULL) != 0, h->alg, Q, q, ov, oi, H, s);
    default: return RMQ_EUNSUPPORTED;
  }
}


// ---- 64-bit form ----------------------------------------------------------------------------
template <int RL>
void launch_chunk_min(const HierView* views, uint32_t chunks, uint32_t clog, float* cv, uint64_t* ci,
                      hipStream_t s) {
  rmq::k_chunk_min<RL><<<(chunks + kBlock - 1) / kBlock, kBlock, 0, s>>>(views, chunks, clog, cv, ci);
}

template <int RL>
void launch_query64(const ulonglong2* Q, uint64_t q, float* ov, uint64_t* oi, const HierView* views,
                    uint32_t clog, uint32_t chunks, const float* cv, const uint64_t* ci, hipStream_t s) {
  const uint32_t grid = static_cast<uint32_t>((q + kBlock - 1) / kBlock);
  if (chunks <= rmq::kLdsChunks)
    rmq::k_query64<RL, true><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, views, clog, chunks, cv, ci);
  else
    rmq::k_query64<RL, false><<<grid, kBlock, 0, s>>>(Q, q, ov, oi, views, clog, chunks, cv, ci);
}

// The sparse table over the top level m = size[top]: floor(log2 m) rows of m entries (row
// stride rounded up to 64). Skipped for the reference's full order, when
// the caller asks, for small top levels and for tables above kTopTableMax entries.
// Shape of the sparse table over the top level m = size[top] (rows = 0: none is built).
struct TopShape {
  uint32_t rows = 0;
  uint64_t stride = 0;
};

TopShape top_shape(const Plan& P, uint32_t flags) {
  TopShape ts;
  const uint64_t m = P.size[P.num_levels - 1];
  if ((flags & (RMQ_BUILD_FULL | RMQ_BUILD_NO_TOP_TABLE)) || m < kTopTableMin) return ts;
  uint32_t rows = 0;
  while ((2ull << rows) <= m) ++rows;  // floor(log2 m)
  const uint64_t stride = (m + 63) & ~63ull;
  if (rows * stride > kTopTableMax) return ts;
  ts.rows = rows;
  ts.stride = stride;
  return ts;
}

// Fills the table (h->tt, already carved out of the hierarchy's memory): one launch for a top
// level of at most kTopSmall entries, else one launch per row.
rmq_status build_top_table(rmq_hierarchy* h, const Plan& P, const TopShape& ts, const float* d_array,
                           hipStream_t s) {
  if (ts.rows == 0) return RMQ_OK;
  const uint32_t top = P.num_levels - 1;
  const uint64_t m = P.size[top];
  const float* tv = top == 0 ? d_array : h->lv + P.off[top];
  const uint8_t* lr = top == 0 ? nullptr : h->lr;  // the top level's offset entries
  const uint32_t roff = h->view.roff[top], sk = P.R_log * top;
  if (m <= rmq::kTopSmall) {
    rmq::k_top_table_small<<<1, rmq::kTopThreads, 0, s>>>(tv, static_cast<uint32_t>(m), h->tt,
                                                          static_cast<uint32_t>(ts.stride), ts.rows, lr, roff, sk);
  } else {
    const uint32_t grid = static_cast<uint32_t>((m + kBlock - 1) / kBlock);
    for (uint32_t j = 1; j <= ts.rows; ++j)
      rmq::k_top_row<<<grid, kBlock, 0, s>>>(tv, static_cast<uint32_t>(m), h->tt, static_cast<uint32_t>(ts.stride), j,
                                             lr, roff, sk);
  }
  h->view.tt = h->tt;
  h->view.tstride = static_cast<uint32_t>(ts.stride);
  return last_error();
}

}  // namespace

void rmq_free(rmq_hierarchy* h);
rmq_status rmq_build_ex(const float* d_array, uint64_t n, uint64_t t, int alg,
                        const rmq_build_options* opt, void* stream, rmq_hierarchy** out);

namespace {

rmq_status build_chunked(const float* d_array, uint64_t n, uint64_t t, int alg, const rmq_build_options* opt,
                         uint32_t clog, const Plan& P, void* stream, rmq_hierarchy** out) {
  if (clog < 8 || clog > kChunkLog) return RMQ_EINVAL;
  rmq_build_options o = opt ? *opt : rmq_build_options{0, 0, 0, 0};
  o.flags &= ~(0xFFu << 8);
  o.flags = (o.flags & ~RMQ_BUILD_BLOCK_SCANS) | RMQ_BUILD_NO_BLOCK_SCANS;  // rmq_query64 reads none
  rmq_hierarchy* h = new (std::nothrow) rmq_hierarchy();
  if (!h) return RMQ_ENOMEM;
  h->alg = alg;
  (void)hipGetDevice(&h->device);
  h->R_log = P.R_log;
  h->group_lanes = P.G;
  h->flags = o.flags;
  h->n = n;
  h->t = t;
  h->buffer_floats = P.buffer_floats;
  h->lv = nullptr;
  h->lr = nullptr;
  h->zeros = nullptr;
  h->view = HierView{};
  h->view.num_levels = P.num_levels;
  for (uint32_t k = 0; k < P.num_levels; ++k) {
    h->msize[k] = P.size[k];
    h->moff[k] = P.off[k];
  }
  h->chunk_log = clog;
  h->chunks = static_cast<uint32_t>((n + (1ull << clog) - 1) >> clog);
  h->part = new (std::nothrow) rmq_hierarchy*[h->chunks]();
  if (!h->part) {
    delete h;
    return RMQ_ENOMEM;
  }
  std::vector<HierView> views(h->chunks);
  rmq_status st = RMQ_OK;
  for (uint32_t c = 0; c < h->chunks && st == RMQ_OK; ++c) {
    const uint64_t b = static_cast<uint64_t>(c) << clog;
    st = rmq_build_ex(d_array + b, std::min<uint64_t>(1ull << clog, n - b), t, alg, &o, stream, &h->part[c]);
    if (st == RMQ_OK) views[c] = h->part[c]->view;
  }
  if (st == RMQ_OK &&
      (hipMalloc(reinterpret_cast<void**>(&h->views), h->chunks * sizeof(HierView)) != hipSuccess ||
       hipMalloc(reinterpret_cast<void**>(&h->cmin), h->chunks * sizeof(float)) != hipSuccess ||
       hipMalloc(reinterpret_cast<void**>(&h->cidx), h->chunks * sizeof(uint64_t)) != hipSuccess)) {
    (void)hipGetLastError();
    st = RMQ_ENOMEM;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (st == RMQ_OK &&
      (hipMemcpyAsync(h->views, views.data(), h->chunks * sizeof(HierView), hipMemcpyHostToDevice, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess))
    st = RMQ_EDEVICE;
  if (st == RMQ_OK) {
    switch (P.R_log) {
      case 2: launch_chunk_min<2>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      case 3: launch_chunk_min<3>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      case 4: launch_chunk_min<4>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      case 5: launch_chunk_min<5>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      case 6: launch_chunk_min<6>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      case 7: launch_chunk_min<7>(h->views, h->chunks, clog, h->cmin, h->cidx, s); break;
      default: st = RMQ_EUNSUPPORTED;
    }
    if (st == RMQ_OK) st = last_error();
  }
  if (st != RMQ_OK) {
    rmq_free(h);
    return st;
  }
  *out = h;
  return RMQ_OK;
}

}  // namespace

extern "C" {


rmq_status rmq_plan(uint64_t n, uint64_t t, int alg, const rmq_build_options* opt,
                    uint32_t* num_levels, uint64_t* sizes, uint64_t* offsets,
                    uint64_t* buffer_floats, uint32_t* R) {
  Plan P;
  const rmq_status st = make_plan(n, t, alg, opt, &P);
  if (st != RMQ_OK) return st;
  if (num_levels) *num_levels = P.num_levels;
  for (uint32_t k = 0; k < P.num_levels; ++k) {
    if (sizes) sizes[k] = P.size[k];
    if (offsets) offsets[k] = P.off[k];
  }
  if (buffer_floats) *buffer_floats = P.buffer_floats;
  if (R) *R = 1u << P.R_log;
  return RMQ_OK;
}

rmq_status rmq_build_ex(const float* d_array, uint64_t n, uint64_t t, int alg,
                        const rmq_build_options* opt, void* stream, rmq_hierarchy** out) {
  if (!out) return RMQ_EINVAL;
  *out = nullptr;
  if (!d_array || (reinterpret_cast<uintptr_t>(d_array) & 15u) != 0) return RMQ_EINVAL;
  Plan P;
  const rmq_status st = make_plan(n, t, alg, opt, &P);
  if (st != RMQ_OK) return st;
  const uint32_t clog = opt ? (opt->flags >> 8) & 0xFFu : 0u;
  if (n > kMaxN || clog != 0) return build_chunked(d_array, n, t, alg, opt, clog ? clog : kChunkLog, P, stream, out);

  rmq_hierarchy* h = new (std::nothrow) rmq_hierarchy();
  if (!h) return RMQ_ENOMEM;
  h->alg = alg;
  if (hipGetDevice(&h->device) != hipSuccess) {
    delete h;
    return RMQ_EDEVICE;
  }
  h->R_log = P.R_log;
  h->group_lanes = P.G;
  h->flags = opt ? opt->flags : 0u;
  h->n = n;
  h->t = t;
  h->buffer_floats = P.buffer_floats;
  // one allocation: the small words (zero signs, level 0's last vector), the level values, the
  // offset entries and the top level's sparse table, each 256-byte aligned
  uint64_t bytes = 0;
  auto carve = [&bytes](uint64_t b) {
    const uint64_t at = bytes;
    bytes += (b + 255) & ~255ull;
    return at;
  };
  // words: [0] zero signs, [4, 8) level 0's last vector, [8] the query check word
  const uint64_t at_words = carve(16 * sizeof(uint32_t));
  uint64_t at_lv = 0, at_lr = 0;
  if (P.num_levels > 1) {
    // padded to the level alignment past the last level too, so that every 16-byte vector that
    // starts inside a level lies inside the allocation
    at_lv = carve(((P.buffer_floats + RMQ_LEVEL_ALIGNMENT - 1) & ~static_cast<uint64_t>(RMQ_LEVEL_ALIGNMENT - 1)) * 4);
    h->rel_bytes = rel_layout(P.R_log, P.num_levels, P.size, h->view.roff);
    at_lr = carve(h->rel_bytes);
  }
  const TopShape ts = top_shape(P, h->flags);
  const uint64_t at_tt = carve(ts.rows * ts.stride * sizeof(uint2));
  // block scans, one 16-byte entry per level-buffer slot: read by the pruned solvers (lane and
  // group) where R >= 16; built on request there, and by default where R >= 32 (tail pieces of
  // up to R - 1 entries); for the levels from the top down while they take at most kScanBytes
  // (a cold level's scans cost more than its vectors)
  const bool pruned = !(h->flags & RMQ_BUILD_FULL);
  uint32_t ps_lo = 0;
  if (pruned && P.num_levels > 1 && P.R_log >= 4 && !(h->flags & RMQ_BUILD_NO_BLOCK_SCANS) &&
      ((h->flags & RMQ_BUILD_BLOCK_SCANS) || P.R_log >= 5)) {
    for (uint32_t k = 1; k < P.num_levels && !ps_lo; ++k)
      if ((P.buffer_floats - P.off[k]) * sizeof(uint4) <= kScanBytes) ps_lo = k;
  }
  const bool scans = ps_lo != 0;
  const uint64_t at_ps = carve(scans ? (P.buffer_floats - P.off[ps_lo]) * sizeof(uint4) : 0);
  const uint64_t at_view = carve(sizeof(HierView));  // the view, on the device, for rmq_query64
  if (hipMalloc(reinterpret_cast<void**>(&h->mem), bytes) != hipSuccess) {
    (void)hipGetLastError();
    delete h;
    return RMQ_ENOMEM;
  }
  h->mem_bytes = bytes;
  h->zeros = reinterpret_cast<uint32_t*>(h->mem + at_words);
  if (P.num_levels > 1) {
    h->lv = reinterpret_cast<float*>(h->mem + at_lv);
    h->lr = h->mem + at_lr;
  }
  if (ts.rows) h->tt = reinterpret_cast<uint2*>(h->mem + at_tt);
  if (scans) {  // indexed from the level buffer's start: ps[off[k] + i] for k >= ps_lo
    h->view.ps = reinterpret_cast<const uint4*>(reinterpret_cast<uintptr_t>(h->mem + at_ps) -
                                                P.off[ps_lo] * sizeof(uint4));
    h->view.ps_lo = ps_lo;
  }
  h->self_view = reinterpret_cast<HierView*>(h->mem + at_view);
  HierView& H = h->view;
  H.A = d_array;
  H.lv = h->lv;
  H.lr = h->lr;
  H.zeros = h->zeros;
  H.tail = reinterpret_cast<const float*>(h->zeros + 4);
  H.num_levels = P.num_levels;
  for (uint32_t k = 0; k < rmq::kMaxLevels; ++k) {
    H.size[k] = k < P.num_levels ? static_cast<uint32_t>(P.size[k]) : 0;
    H.off[k] = k < P.num_levels ? static_cast<uint32_t>(P.off[k]) : 0;
  }
  const hipStream_t s = static_cast<hipStream_t>(stream);
  // the level-0 pass records the zero signs and NaN; with no such pass all are assumed
  rmq::k_init_words<<<1, 1, 0, s>>>(h->zeros, P.num_levels > 1 ? 0u : rmq::kLevel0Unknown, d_array,
                                    static_cast<uint32_t>(n));
  if (last_error() != RMQ_OK) {
    rmq_free(h);
    return RMQ_EDEVICE;
  }
  const rmq_status bs = build_levels(P.R_log, H, s);
  if (bs != RMQ_OK) {
    rmq_free(h);
    return bs;
  }
  const rmq_status tst = build_top_table(h, P, ts, d_array, s);
  if (tst != RMQ_OK) {
    rmq_free(h);
    return tst;
  }
  if (scans && build_block_scans(P.R_log, H, s) != RMQ_OK) {
    rmq_free(h);
    return RMQ_EDEVICE;
  }
  // the finished view (h lives until rmq_free, so the copy's source outlives it)
  if (hipMemcpyAsync(h->self_view, &h->view, sizeof(HierView), hipMemcpyHostToDevice, s) != hipSuccess) {
    (void)hipGetLastError();
    rmq_free(h);
    return RMQ_EDEVICE;
  }
  *out = h;
  return RMQ_OK;
}

rmq_status rmq_build(const float* d_array, uint64_t n, uint64_t t, int alg, void* stream,
                     rmq_hierarchy** out) {
  return rmq_build_ex(d_array, n, t, alg, nullptr, stream, out);
}

rmq_status rmq_query(const rmq_hierarchy* h, const uint32_t* d_lr, uint64_t q, float* d_val,
                     uint32_t* d_idx, void* stream) {
  if (!h || h->chunk_log) return RMQ_EINVAL;  // a chunked (64-bit) hierarchy: rmq_query64
  if (q == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx)) return RMQ_EINVAL;
  if ((reinterpret_cast<uintptr_t>(d_lr) & 7u) != 0) return RMQ_EINVAL;
  if (q > 0xFFFFFFFFull) return RMQ_EINVAL;
  if (h->flags & RMQ_BUILD_CHECK_QUERIES) {
    const rmq_status cs = check_queries(h, reinterpret_cast<const uint2*>(d_lr), q, static_cast<hipStream_t>(stream));
    if (cs != RMQ_OK) return cs;
  }
  // one launch per kLaunchQueries queries (the one-lane-per-query kernels have no grid stride)
  const uint2* Q = reinterpret_cast<const uint2*>(d_lr);
  for (uint64_t q0 = 0; q0 < q; q0 += kLaunchQueries) {
    const uint32_t cnt = static_cast<uint32_t>(std::min<uint64_t>(kLaunchQueries, q - q0));
    const rmq_status st = query_dispatch(h, Q + q0, cnt, d_val ? d_val + q0 : nullptr, d_idx ? d_idx + q0 : nullptr,
                                         static_cast<hipStream_t>(stream));
    if (st != RMQ_OK) return st;
  }
  return RMQ_OK;
}

void rmq_free(rmq_hierarchy* h) {
  if (!h) return;
  for (uint32_t c = 0; h->part && c < h->chunks; ++c) rmq_free(h->part[c]);
  delete[] h->part;
  if (h->views) (void)hipFree(h->views);
  if (h->cmin) (void)hipFree(h->cmin);
  if (h->cidx) (void)hipFree(h->cidx);
  if (h->mem) (void)hipFree(h->mem);
  delete h;
}

rmq_status rmq_get_metadata(const rmq_hierarchy* h, uint32_t* num_levels, uint64_t* sizes,
                            uint64_t* offsets, uint64_t* buffer_floats, uint32_t* R) {
  if (!h) return RMQ_EINVAL;
  const HierView& H = h->view;
  if (num_levels) *num_levels = H.num_levels;
  for (uint32_t k = 0; k < H.num_levels; ++k) {
    if (sizes) sizes[k] = h->chunk_log ? h->msize[k] : H.size[k];
    if (offsets) offsets[k] = k == 0 ? 0xFFFFFFFFull : (h->chunk_log ? h->moff[k] : H.off[k]);
  }
  if (buffer_floats) *buffer_floats = h->buffer_floats;
  if (R) *R = 1u << h->R_log;
  return RMQ_OK;
}

rmq_status rmq_read_levels(const rmq_hierarchy* h, float* d_val, uint32_t* d_idx, void* stream) {
  if (!h || (!d_val && !d_idx)) return RMQ_EINVAL;
  if (h->chunk_log) return RMQ_EUNSUPPORTED;  // the 64-bit form keeps its levels per chunk
  return rmq_internal_read_levels(h->view, h->R_log, d_val, d_idx, stream);
}

rmq_status rmq_build_each(const float* const* d_arrays, const uint64_t* n, const uint64_t* t,
                          int alg, const rmq_build_options* opt, uint32_t count, void* stream,
                          rmq_hierarchy** out) {
  if (!d_arrays || !n || !t || !out) return RMQ_EINVAL;
  for (uint32_t i = 0; i < count; ++i) {
    const rmq_status st = rmq_build_ex(d_arrays[i], n[i], t[i], alg, opt, stream, &out[i]);
    if (st != RMQ_OK) {
      for (uint32_t j = 0; j < i; ++j) {
        rmq_free(out[j]);
        out[j] = nullptr;
      }
      return st;
    }
  }
  return RMQ_OK;
}

rmq_status rmq_query_each(rmq_hierarchy* const* hs, const uint32_t* const* d_lr, const uint64_t* q,
                          float* const* d_val, uint32_t* const* d_idx, uint32_t count, void* stream) {
  if (!hs || !d_lr || !q) return RMQ_EINVAL;
  for (uint32_t i = 0; i < count; ++i) {
    const rmq_status st = rmq_query(hs[i], d_lr[i], q[i], d_val ? d_val[i] : nullptr,
                                    d_idx ? d_idx[i] : nullptr, stream);
    if (st != RMQ_OK) return st;
  }
  return RMQ_OK;
}

rmq_status rmq_gpu_base(const float* d_array, uint64_t n, const uint32_t* d_lr, uint64_t q,
                        float* d_val, uint32_t* d_idx, void* stream) {
  if (!d_array || n == 0 || n > kMaxN) return RMQ_EINVAL;
  if (q == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx) || q > 0xFFFFFFFFull) return RMQ_EINVAL;
  if ((reinterpret_cast<uintptr_t>(d_lr) & 7u) != 0) return RMQ_EINVAL;
  const uint2* Q = reinterpret_cast<const uint2*>(d_lr);
  for (uint64_t q0 = 0; q0 < q; q0 += kLaunchQueries) {
    const uint64_t cnt = std::min<uint64_t>(kLaunchQueries, q - q0);
    rmq::k_gpu_base<<<grid_for(cnt), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
        d_array, Q + q0, static_cast<uint32_t>(cnt), d_val ? d_val + q0 : nullptr, d_idx ? d_idx + q0 : nullptr);
    if (last_error() != RMQ_OK) return RMQ_EDEVICE;
  }
  return RMQ_OK;
}

rmq_status rmq_query64(const rmq_hierarchy* h, const uint64_t* d_lr, uint64_t q, float* d_val,
                       uint64_t* d_idx, void* stream) {
  if (!h) return RMQ_EINVAL;
  if (q == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx) || (reinterpret_cast<uintptr_t>(d_lr) & 15u) != 0) return RMQ_EINVAL;
  if ((q + kBlock - 1) / kBlock > 0xFFFFFFFFull) return RMQ_EINVAL;
  const hipStream_t s = static_cast<hipStream_t>(stream);
  if (h->flags & RMQ_BUILD_CHECK_QUERIES) {
    const rmq_status cs = check_queries(h, reinterpret_cast<const ulonglong2*>(d_lr), q, s);
    if (cs != RMQ_OK) return cs;
  }
  // a one-piece hierarchy is its own single chunk: every position is in chunk 0
  const uint32_t clog = h->chunk_log ? h->chunk_log : 63u;
  const uint32_t chunks = h->chunk_log ? h->chunks : 1u;
  const HierView* views = h->chunk_log ? h->views : h->self_view;
  const float* cv = h->cmin;
  const uint64_t* ci = h->cidx;
  const ulonglong2* Q = reinterpret_cast<const ulonglong2*>(d_lr);
  switch (h->R_log) {
    case 2: launch_query64<2>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    case 3: launch_query64<3>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    case 4: launch_query64<4>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    case 5: launch_query64<5>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    case 6: launch_query64<6>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    case 7: launch_query64<7>(Q, q, d_val, d_idx, views, clog, chunks, cv, ci, s); break;
    default: return RMQ_EUNSUPPORTED;
  }
  return last_error();
}

rmq_status rmq_gpu_base64(const float* d_array, uint64_t n, const uint64_t* d_lr, uint64_t q, float* d_val,
                          uint64_t* d_idx, void* stream) {
  if (!d_array || n == 0 || n > kMaxN64) return RMQ_EINVAL;
  if (q == 0) return RMQ_OK;
  if (!d_lr || (!d_val && !d_idx) || (reinterpret_cast<uintptr_t>(d_lr) & 15u) != 0) return RMQ_EINVAL;
  if ((q + kBlock - 1) / kBlock > 0xFFFFFFFFull) return RMQ_EINVAL;
  rmq::k_gpu_base64<<<static_cast<uint32_t>((q + kBlock - 1) / kBlock), kBlock, 0, static_cast<hipStream_t>(stream)>>>(
      d_array, reinterpret_cast<const ulonglong2*>(d_lr), q, d_val, d_idx);
  return last_error();
}

const char* rmq_status_str(rmq_status s) {
  switch (s) {
    case RMQ_OK: return "ok";
    case RMQ_EINVAL: return "invalid argument";
    case RMQ_ETHRESHOLD: return "scan threshold too small for the reduction factor";
    case RMQ_ENOMEM: return "out of device memory";
    case RMQ_EDEVICE: return "HIP runtime error";
    case RMQ_EUNSUPPORTED: return "unsupported option";
    case RMQ_EAGAIN: return "staging full, poll and submit again";
  }
  return "unknown status";
}

const char* rmq_version(void) { return "rmq-mi355x 0.1 gfx950"; }

}  // extern "C"

const rmq::HierView* rmq_internal_view(const rmq_hierarchy* h) { return h ? &h->view : nullptr; }

uint64_t rmq_internal_level_bytes(const rmq_hierarchy* h) { return h ? h->mem_bytes : 0; }

rmq_status rmq_internal_read_levels(const HierView& H, uint32_t R_log, float* d_val, uint32_t* d_idx,
                                    void* stream) {
  const hipStream_t s = static_cast<hipStream_t>(stream);
  for (uint32_t k = 1; k < H.num_levels; ++k) {
    rmq::k_read_level<<<grid_for(H.size[k]), kBlock, 0, s>>>(H, k, R_log * k, d_val, d_idx);
    if (last_error() != RMQ_OK) return RMQ_EDEVICE;
  }
  return RMQ_OK;
}

uint64_t rmq_internal_rel_layout(uint32_t R_log, uint32_t num_levels, const uint64_t* size, uint32_t* roff) {
  return rel_layout(R_log, num_levels, size, roff);
}
