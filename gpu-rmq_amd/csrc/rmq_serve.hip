// rmq_serve.hip — serving loop over device-resident hierarchies (include/rmq_serve.h).
//
// Resident arrays are ordinary hierarchies (rmq_build_ex; R = 4, 8 or 32 by variant) whose level
// views sit in a device table; a batch is a run of pending requests, one request = one segment of
// tiles on its array, answered by the multi-array persistent kernel (rmq_multi.hpp). Requests are
// staged once: submit checks and copies their pairs into a ring of pinned, device-coherent host
// memory, which the kernel reads in place, and the kernel writes the answers into a second ring
// at the same positions, from which retiring copies them to the caller. A batch is therefore two
// kernel launches and an event, with no copy commands and no per-request allocation; up to
// `slots` batches overlap on the device while the host stages and packs the next ones.

#include <hip/hip_runtime.h>
#include <emmintrin.h>

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <deque>
#include <new>
#include <vector>

#include "../../include/rmq_serve.h"
#include "rmq_internal.hpp"
#include "rmq_multi.hpp"

namespace {

struct Resident {
  bool live = false;
  float* owned = nullptr;  // device array owned by the server (HOST / UNIFORM admission)
  rmq_hierarchy* h = nullptr;
  uint64_t n = 0;          // elements: submit refuses pairs outside [0, n)
  uint64_t bytes = 0;
  uint32_t users = 0;      // pending + in-flight requests
};

// A request from submit until it is reported done. Its pairs (and then its answers) live at ring
// positions [pos, pos + q) (absolute; the ring index is pos % ring_cap, and a request never wraps).
struct Request {
  uint64_t ticket = 0;
  int32_t array = -1;
  uint32_t q = 0;
  uint64_t pos = 0;
  float* h_val = nullptr;
  uint32_t* h_idx = nullptr;
  bool done = false;
};

struct Slot {
  hipStream_t stream = nullptr;
  hipEvent_t done = nullptr;
  bool busy = false;
  uint64_t t0 = 0, t1 = 0;  // tickets [t0, t1) of the batch
  // pinned host memory, coherent with the device; d_* are the device's addresses of the same
  rmq::QueryTile* h_tiles = nullptr;
  uint32_t* h_qstart = nullptr;
  rmq::QueryTile* d_tiles = nullptr;
  uint32_t* d_qstart = nullptr;
  uint32_t* d_counter = nullptr;  // device memory
};

bool dalloc(void** p, size_t bytes) {
  if (hipMalloc(p, bytes ? bytes : 4) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return false;
  }
  return true;
}

bool halloc(void** p, size_t bytes) {
  if (hipHostMalloc(p, bytes ? bytes : 4, hipHostMallocDefault) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return false;
  }
  return true;
}

// Pinned host memory that kernels read and write in place (coherent: no GPU caching, so the
// host's writes before a launch and the kernel's writes before its completion are seen).
template <typename T>
bool zalloc(T** h, T** d, size_t bytes) {
  if (hipHostMalloc(reinterpret_cast<void**>(h), bytes ? bytes : 4, hipHostMallocMapped | hipHostMallocCoherent) !=
      hipSuccess) {
    (void)hipGetLastError();
    *h = nullptr;
    return false;
  }
  if (hipHostGetDevicePointer(reinterpret_cast<void**>(d), *h, 0) != hipSuccess) {
    (void)hipGetLastError();
    return false;
  }
  return true;
}

}  // namespace

struct rmq_server {
  rmq_server_options opt{};
  uint32_t max_tiles = 0;
  std::vector<Resident> arrays;
  rmq::HierView* d_views = nullptr;
  uint32_t* d_variant = nullptr;
  rmq::HierView* h_views = nullptr;  // pinned mirrors of the tables
  uint32_t* h_variant = nullptr;
  hipStream_t admin = nullptr;
  // the staging rings: pairs, value answers, index answers (ring_cap entries each)
  uint64_t ring_cap = 0;
  uint32_t* h_ring_q = nullptr;
  float* h_ring_v = nullptr;
  uint32_t* h_ring_i = nullptr;
  uint32_t* d_ring_q = nullptr;
  float* d_ring_v = nullptr;
  uint32_t* d_ring_i = nullptr;
  uint64_t head = 0;  // next free ring position (absolute)
  // every request not yet reported done, in ticket order: the first `launched` are in batches
  std::deque<Request> live;
  size_t launched = 0;
  std::vector<Slot> slots;
  uint64_t next_ticket = 1;
  rmq_server_stats stats{};
  // per-batch scratch of the packer (kept to avoid reallocating)
  std::vector<uint64_t> seg_start, seg_cnt;
  std::vector<uint32_t> seg_ids;
  std::vector<rmq::QueryTile> tiles;
  std::vector<uint32_t> qstart;
};

namespace {

void free_slot(Slot& sl) {
  if (sl.stream) (void)hipStreamSynchronize(sl.stream);
  (void)hipHostFree(sl.h_tiles);
  (void)hipHostFree(sl.h_qstart);
  (void)hipFree(sl.d_counter);
  if (sl.done) (void)hipEventDestroy(sl.done);
  if (sl.stream) (void)hipStreamDestroy(sl.stream);
  sl = Slot();
}

bool init_slot(Slot& sl, uint32_t max_tiles) {
  return hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking) == hipSuccess &&
         hipEventCreateWithFlags(&sl.done, hipEventDisableTiming) == hipSuccess &&
         zalloc(&sl.h_tiles, &sl.d_tiles, max_tiles * sizeof(rmq::QueryTile)) &&
         zalloc(&sl.h_qstart, &sl.d_qstart, (rmq::kQueues + 1) * 4) &&
         dalloc(reinterpret_cast<void**>(&sl.d_counter), rmq::kCounterWords * 4);
}

// The oldest ring position still held by a request (head when none is).
uint64_t ring_tail(const rmq_server* s) { return s->live.empty() ? s->head : s->live.front().pos; }

// Copy a finished slot's answers to its requests' buffers, mark them done and free the slot; the
// requests at the front of the queue that are done release their ring space.
void retire(rmq_server* s, Slot& sl, uint64_t* done, uint32_t max_done, uint32_t* n_done) {
  const uint64_t first = s->live.front().ticket;
  for (uint64_t t = sl.t0; t < sl.t1; ++t) {
    Request& r = s->live[t - first];
    const uint64_t at = r.pos % s->ring_cap;
    if (r.h_val) std::memcpy(r.h_val, s->h_ring_v + at, r.q * 4ull);
    if (r.h_idx) std::memcpy(r.h_idx, s->h_ring_i + at, r.q * 4ull);
    r.done = true;
    s->arrays[r.array].users--;
    s->stats.completed++;
    if (done && *n_done < max_done) done[(*n_done)++] = r.ticket;
  }
  sl.busy = false;
  s->stats.in_flight--;
  while (!s->live.empty() && s->live.front().done) {
    s->live.pop_front();
    s->launched--;
  }
}

// Pack the oldest pending requests into a free slot and launch the batch.
rmq_status launch(rmq_server* s, Slot& sl) {
  const size_t p0 = s->launched;
  uint64_t total = 0;
  bool want_idx = false, lanes_only = true;
  s->seg_start.clear();
  s->seg_cnt.clear();
  s->seg_ids.clear();
  size_t p = p0;
  while (p < s->live.size() && p - p0 < s->opt.max_batch_requests && total + s->live[p].q <= s->opt.max_batch_queries) {
    const Request& r = s->live[p];
    s->seg_start.push_back(r.pos % s->ring_cap);
    s->seg_cnt.push_back(r.q);
    s->seg_ids.push_back(static_cast<uint32_t>(r.array));
    want_idx = want_idx || r.h_idx != nullptr;
    lanes_only = lanes_only && rmq::variant_is_lane(s->h_variant[r.array]);
    total += r.q;
    ++p;
  }
  if (p == p0) return RMQ_OK;
  // within kMaxTiles: at most max_batch_queries / 256 + max_batch_requests tiles
  (void)rmq::make_segment_tiles(static_cast<uint32_t>(s->seg_ids.size()), s->seg_start.data(), s->seg_cnt.data(),
                                s->tiles, s->qstart, s->seg_ids.data());
  std::memcpy(sl.h_tiles, s->tiles.data(), s->tiles.size() * sizeof(rmq::QueryTile));
  std::memcpy(sl.h_qstart, s->qstart.data(), s->qstart.size() * 4);
  rmq::launch_multi_query(s->d_views, s->d_variant, sl.d_tiles, sl.d_qstart, sl.d_counter,
                          static_cast<uint32_t>(s->tiles.size()), reinterpret_cast<const uint2*>(s->d_ring_q),
                          s->d_ring_v, want_idx ? s->d_ring_i : nullptr, sl.stream, lanes_only);
  bool ok = hipGetLastError() == hipSuccess;
  ok = ok && hipEventRecord(sl.done, sl.stream) == hipSuccess;
  sl.t0 = s->live[p0].ticket;
  sl.t1 = sl.t0 + (p - p0);
  sl.busy = true;
  s->launched = p;
  s->stats.in_flight++;
  s->stats.batches++;
  s->stats.queries += total;
  return ok ? RMQ_OK : RMQ_EDEVICE;
}

// Copies q {l, r} pairs from src into the staging ring at dst and returns nonzero when any pair
// is outside 0 <= l <= r < n (n < 2^31). Two pairs per 16-byte step, checked with signed
// compares of the values offset by 2^31 (which keeps their unsigned order), and written with
// streaming stores: the ring is written once here and read by the device, so its lines are not
// read into the cache first. The fence orders those stores before the launch that reads them.
uint32_t copy_checked(uint32_t* dst, const uint32_t* src, uint64_t q, uint64_t n) {
  uint32_t bad = 0;
  uint64_t i = 0;
  auto one = [&](uint64_t k) {
    const uint32_t l = src[2 * k], r = src[2 * k + 1];
    bad |= static_cast<uint32_t>(l > r) | static_cast<uint32_t>(r >= n);
    dst[2 * k] = l;
    dst[2 * k + 1] = r;
  };
  if ((reinterpret_cast<uintptr_t>(dst) & 15u) != 0 && q > 0) one(i++);  // pairs are 8-byte aligned
  const __m128i bias = _mm_set1_epi32(static_cast<int32_t>(0x80000000u));
  const __m128i nb = _mm_set1_epi32(static_cast<int32_t>(static_cast<uint32_t>(n) ^ 0x80000000u));
  const __m128i lanes_l = _mm_set_epi32(0, -1, 0, -1);  // elements 0 and 2: l; 1 and 3: r
  __m128i acc = _mm_setzero_si128();
  for (; i + 2 <= q; i += 2) {
    const __m128i v = _mm_loadu_si128(reinterpret_cast<const __m128i*>(src + 2 * i));
    const __m128i b = _mm_xor_si128(v, bias);
    const __m128i sw = _mm_shuffle_epi32(b, _MM_SHUFFLE(2, 3, 0, 1));  // r0 l0 r1 l1
    const __m128i l_gt_r = _mm_and_si128(_mm_cmpgt_epi32(b, sw), lanes_l);
    const __m128i r_ge_n = _mm_andnot_si128(_mm_or_si128(_mm_cmpgt_epi32(nb, b), lanes_l), _mm_set1_epi32(-1));
    acc = _mm_or_si128(acc, _mm_or_si128(l_gt_r, r_ge_n));
    _mm_stream_si128(reinterpret_cast<__m128i*>(dst + 2 * i), v);
  }
  if (i < q) one(i);
  _mm_sfence();
  return bad | static_cast<uint32_t>(_mm_movemask_epi8(acc) != 0);
}

}  // namespace

extern "C" {

rmq_status rmq_server_create(const rmq_server_options* opt, rmq_server** out) {
  if (!out) return RMQ_EINVAL;
  *out = nullptr;
  rmq_server* s = new (std::nothrow) rmq_server();
  if (!s) return RMQ_ENOMEM;
  if (opt) s->opt = *opt;
  if (!s->opt.slots) s->opt.slots = 3;
  if (!s->opt.max_batch_queries) s->opt.max_batch_queries = 1u << 20;
  if (!s->opt.reserve_bytes) s->opt.reserve_bytes = 2ull << 30;
  if (!s->opt.max_arrays) s->opt.max_arrays = 1024;
  if (!s->opt.max_batch_requests) s->opt.max_batch_requests = 4096;
  s->max_tiles = s->opt.max_batch_queries / rmq::kTileQueries + s->opt.max_batch_requests;
  if (!s->opt.ring_queries)
    s->opt.ring_queries = std::max<uint64_t>(1ull << 22, 2ull * s->opt.slots * s->opt.max_batch_queries);
  if (s->opt.max_batch_requests > rmq::kMaxTiles / 2 || s->opt.ring_queries < s->opt.max_batch_queries ||
      s->opt.ring_queries > (1ull << 32)) {
    delete s;
    return RMQ_EINVAL;
  }
  s->ring_cap = s->opt.ring_queries;
  bool ok = hipStreamCreateWithFlags(&s->admin, hipStreamNonBlocking) == hipSuccess &&
            dalloc(reinterpret_cast<void**>(&s->d_views), s->opt.max_arrays * sizeof(rmq::HierView)) &&
            dalloc(reinterpret_cast<void**>(&s->d_variant), s->opt.max_arrays * 4) &&
            halloc(reinterpret_cast<void**>(&s->h_views), s->opt.max_arrays * sizeof(rmq::HierView)) &&
            halloc(reinterpret_cast<void**>(&s->h_variant), s->opt.max_arrays * 4) &&
            zalloc(&s->h_ring_q, &s->d_ring_q, s->ring_cap * 8) && zalloc(&s->h_ring_v, &s->d_ring_v, s->ring_cap * 4) &&
            zalloc(&s->h_ring_i, &s->d_ring_i, s->ring_cap * 4);
  s->slots.resize(s->opt.slots);
  for (Slot& sl : s->slots) ok = ok && init_slot(sl, s->max_tiles);
  if (!ok) {
    rmq_server_destroy(s);
    return RMQ_ENOMEM;
  }
  *out = s;
  return RMQ_OK;
}

void rmq_server_destroy(rmq_server* s) {
  if (!s) return;
  for (Slot& sl : s->slots) free_slot(sl);
  for (Resident& r : s->arrays) {
    if (r.h) rmq_free(r.h);
    if (r.owned) (void)hipFree(r.owned);
  }
  (void)hipFree(s->d_views);
  (void)hipFree(s->d_variant);
  (void)hipHostFree(s->h_views);
  (void)hipHostFree(s->h_variant);
  (void)hipHostFree(s->h_ring_q);
  (void)hipHostFree(s->h_ring_v);
  (void)hipHostFree(s->h_ring_i);
  if (s->admin) (void)hipStreamDestroy(s->admin);
  delete s;
}

rmq_status rmq_server_admit(rmq_server* s, int source, const float* data, uint64_t n, uint64_t seed,
                            uint32_t t, uint32_t variant, int32_t* array_id) {
  if (!s || !array_id || n == 0 || n > 0x7FFFFFFFull || !rmq::variant_ok(variant)) return RMQ_EINVAL;
  if (source != RMQ_ADMIT_DEVICE && source != RMQ_ADMIT_HOST && source != RMQ_ADMIT_UNIFORM)
    return RMQ_EINVAL;
  if (source != RMQ_ADMIT_UNIFORM && !data) return RMQ_EINVAL;
  int alg = 0;
  rmq_build_options o;
  rmq::variant_build(variant, &alg, &o);
  uint64_t floats = 0;
  rmq_status st = rmq_plan(n, t, alg, &o, nullptr, nullptr, nullptr, &floats, nullptr);
  if (st != RMQ_OK) return st;
  const uint64_t need = floats * 8 + (source == RMQ_ADMIT_DEVICE ? 0 : n * 4) + (1u << 16);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess) return RMQ_EDEVICE;
  if (free_b < need + s->opt.reserve_bytes) return RMQ_ENOMEM;
  int32_t id = -1;
  for (size_t i = 0; i < s->arrays.size(); ++i)
    if (!s->arrays[i].live) {
      id = static_cast<int32_t>(i);
      break;
    }
  if (id < 0) {
    if (s->arrays.size() >= s->opt.max_arrays) return RMQ_ENOMEM;
    s->arrays.emplace_back();
    id = static_cast<int32_t>(s->arrays.size() - 1);
  }
  Resident& r = s->arrays[id];
  const float* A = data;
  // a borrowed device array may still be written by work on other streams (the caller's fill or
  // copy): the build below runs on the server's own non-blocking stream, so wait for the device
  if (source == RMQ_ADMIT_DEVICE && hipDeviceSynchronize() != hipSuccess) {
    (void)hipGetLastError();
    return RMQ_EDEVICE;
  }
  if (source != RMQ_ADMIT_DEVICE) {
    if (!dalloc(reinterpret_cast<void**>(&r.owned), n * 4)) return RMQ_ENOMEM;
    if (source == RMQ_ADMIT_HOST) {
      if (hipMemcpyAsync(r.owned, data, n * 4, hipMemcpyHostToDevice, s->admin) != hipSuccess) st = RMQ_EDEVICE;
    } else {
      st = rmq_fill_uniform(r.owned, n, seed, 0, s->admin);
    }
    A = r.owned;
  }
  if (st == RMQ_OK) st = rmq_build_ex(A, n, t, alg, &o, s->admin, &r.h);
  if (st == RMQ_OK) {
    s->h_views[id] = *rmq_internal_view(r.h);
    s->h_variant[id] = variant;
    if (hipMemcpyAsync(s->d_views + id, s->h_views + id, sizeof(rmq::HierView), hipMemcpyHostToDevice,
                       s->admin) != hipSuccess ||
        hipMemcpyAsync(s->d_variant + id, s->h_variant + id, 4, hipMemcpyHostToDevice, s->admin) != hipSuccess ||
        hipStreamSynchronize(s->admin) != hipSuccess)
      st = RMQ_EDEVICE;
  }
  if (st != RMQ_OK) {
    if (r.h) rmq_free(r.h);
    if (r.owned) (void)hipFree(r.owned);
    r = Resident();
    return st;
  }
  r.live = true;
  r.n = n;
  r.bytes = rmq_internal_level_bytes(r.h) + (r.owned ? n * 4 : 0);
  s->stats.resident++;
  s->stats.resident_bytes += r.bytes;
  *array_id = id;
  return RMQ_OK;
}

rmq_status rmq_server_evict(rmq_server* s, int32_t array_id) {
  if (!s || array_id < 0 || static_cast<size_t>(array_id) >= s->arrays.size()) return RMQ_EINVAL;
  Resident& r = s->arrays[array_id];
  if (!r.live || r.users) return RMQ_EINVAL;
  rmq_free(r.h);
  if (r.owned) (void)hipFree(r.owned);
  s->stats.resident--;
  s->stats.resident_bytes -= r.bytes;
  r = Resident();
  return RMQ_OK;
}

rmq_status rmq_server_submit(rmq_server* s, int32_t array_id, const uint32_t* h_lr, uint64_t q,
                             float* h_val, uint32_t* h_idx, uint64_t* ticket) {
  if (!s || !h_lr || !ticket || q == 0 || q > s->opt.max_batch_queries || (!h_val && !h_idx))
    return RMQ_EINVAL;
  if (array_id < 0 || static_cast<size_t>(array_id) >= s->arrays.size() || !s->arrays[array_id].live)
    return RMQ_EINVAL;
  // ring space: q consecutive entries, starting over at the ring's beginning when they would
  // not fit before its end. An empty ring always has room (q <= max_batch_queries <= ring_cap):
  // the space skipped at its end then holds nothing, so it must not count against the request.
  uint64_t pos = s->head;
  if (pos % s->ring_cap + q > s->ring_cap) pos += s->ring_cap - pos % s->ring_cap;
  if (!s->live.empty() && pos + q - ring_tail(s) > s->ring_cap) return RMQ_EAGAIN;
  // every pair is checked while it is copied: one bad pair refuses the whole request before
  // anything is queued, so no kernel ever reads outside a resident array
  const uint64_t n = s->arrays[array_id].n;
  if (copy_checked(s->h_ring_q + 2 * (pos % s->ring_cap), h_lr, q, n)) return RMQ_EINVAL;
  Request r;
  r.ticket = s->next_ticket++;
  r.array = array_id;
  r.q = static_cast<uint32_t>(q);
  r.pos = pos;
  r.h_val = h_val;
  r.h_idx = h_idx;
  s->head = pos + q;
  s->arrays[array_id].users++;
  s->stats.submitted++;
  *ticket = r.ticket;
  s->live.push_back(r);
  return RMQ_OK;
}

rmq_status rmq_server_poll(rmq_server* s, uint64_t* done, uint32_t max_done, uint32_t* n_done) {
  if (!s) return RMQ_EINVAL;
  uint32_t nd = 0;
  rmq_status st = RMQ_OK;
  // batches finish in any order; ring space is released in ticket order (retire)
  for (Slot& sl : s->slots) {
    if (!sl.busy) continue;
    const hipError_t e = hipEventQuery(sl.done);
    if (e == hipSuccess) {
      retire(s, sl, done, max_done, &nd);
    } else if (e != hipErrorNotReady) {
      st = RMQ_EDEVICE;
    }
  }
  for (Slot& sl : s->slots) {
    if (s->launched == s->live.size() || st != RMQ_OK) break;
    if (!sl.busy) st = launch(s, sl);
  }
  if (n_done) *n_done = nd;
  return st;
}

rmq_status rmq_server_drain(rmq_server* s, uint64_t* done, uint32_t max_done, uint32_t* n_done) {
  if (!s) return RMQ_EINVAL;
  uint32_t total = 0;
  while (true) {
    uint32_t nd = 0;
    const rmq_status st = rmq_server_poll(s, done ? done + total : nullptr,
                                          max_done > total ? max_done - total : 0, &nd);
    total += nd;
    if (st != RMQ_OK) {
      if (n_done) *n_done = total;
      return st;
    }
    Slot* wait = nullptr;
    for (Slot& sl : s->slots)
      if (sl.busy) {
        wait = &sl;
        break;
      }
    if (!wait && s->launched == s->live.size()) break;
    if (wait && hipEventSynchronize(wait->done) != hipSuccess) {
      if (n_done) *n_done = total;
      return RMQ_EDEVICE;
    }
  }
  if (n_done) *n_done = total;
  return RMQ_OK;
}

rmq_status rmq_server_get_stats(const rmq_server* s, rmq_server_stats* st) {
  if (!s || !st) return RMQ_EINVAL;
  *st = s->stats;
  return RMQ_OK;
}

}  // extern "C"
