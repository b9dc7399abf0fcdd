// rtxrmq_algs.hpp — host mirror of the reference's per-algorithm entry points, on top of the
// C ABI. Same names, argument meaning and ownership as the reference:
//   float* xxx_rmq(n, t, q, d_array, d_query, args)                  src/xxx.h:203-298
//   float* xxx_rmq_multi_load(n, t, q, d_array, d_query, args)       src/xxx.h:530-627
//   float* hierarchical_vector_load(n, t, q, d_array, d_query, args) src/xxx.h:932-1040
// Each builds the hierarchy, answers the queries `reps` times (host wall clock around
// launch + synchronise, as the reference times it), prints the reference's progress lines,
// appends the CSV fields, and returns a malloc'd host array the caller frees. An invalid t
// prints the reference's message, completes the CSV line and exits(EXIT_FAILURE).
// The *_idx variants return the leftmost argmin indices (no reference counterpart; they follow
// CPU BASE IDX, src/cpu_methods.h:42-53). Every entry point is a template on the position type
// P, the reference's nType (src/main.cu:21-27): uint32_t, or uint64_t for its IS_LONG build
// (64-bit {l, r} pairs and indices, rmq_query64 / rmq_gpu_base64).
// cpu_rmq / cpu_rmq_idx (algs 0 / 100) are the reference's CPU BASE (src/cpu_methods.h:18-39,
// 55-77): a scan per query on the host, OpenMP over queries with --nt threads.
#pragma once

#include <hip/hip_runtime_api.h>
#include <omp.h>
#include <unistd.h>

#include <chrono>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <ctime>
#include <string>

#include "../../include/rmq.h"

namespace rtxrmq {

struct CmdArgs {  // src/tools.h:26-35
  int64_t n = 0, nb = 0;
  uint32_t q = 0, log_bs = 15, log_build_threshold = 0;
  int lr = 0, alg = 0, reps = 10, dev = 0, nt = 1, seed = 0;
  int check = 0, trivialCheck = 0, save_input_data = 0, randTrivialCheck = 0;
  int save_time = 0, save_power = 0;
  uint32_t cg_size_log = 3, cg_amount_log = 2, group_lanes = 0;  // runtime forms of the
                                                                 // compile-time constants
  int is_long = 0;  // IS_LONG (CMakeLists.txt:57-59): --long, or n > 2^31 - 1
  int pack_descents = 0;  // RMQ_BUILD_PACK_DESCENTS
  std::string time_file, power_file;
};

struct VBHMem {
  size_t out_buffer = 0;
  size_t temp_buffer = 0;
};

class Timer {  // common/Timer.h:36-78
 public:
  Timer() { restart(); }
  void restart() { t0_ = clock::now(); }
  double ms() const {
    return std::chrono::duration<double, std::milli>(clock::now() - t0_).count();
  }

 private:
  using clock = std::chrono::high_resolution_clock;
  clock::time_point t0_;
};

inline const char* csv_path(const CmdArgs& a) {
  return a.time_file.empty() ? "../results/data.csv" : a.time_file.c_str();
}

// complete_line, src/device_tools.cuh:20-26
inline void complete_line(const CmdArgs& a) {
  if (!a.save_time) return;
  if (FILE* fp = fopen(csv_path(a), "a")) {
    fprintf(fp, "\n");
    fclose(fp);
  }
}

// write_results(time, q, construction, reps, args, mem), src/tools.h:455-486
inline void write_results(double time_ms, uint32_t q, double construction_ms, int reps,
                          const CmdArgs& a, VBHMem mem) {
  if (!a.save_time) return;
  size_t free_b = 0, total_b = 0;
  const int ok = hipMemGetInfo(&free_b, &total_b) == hipSuccess ? 1 : 0;
  const double time_it = time_ms / reps;
  if (FILE* fp = fopen(csv_path(a), "a")) {
    fprintf(fp, "%f,%f,%f,%f,%f,%f,%i,%f", time_ms / 1000.0, (double)q / (time_it / 1000.0),
            time_it * 1e6 / q, construction_ms, mem.out_buffer / 1e6, mem.temp_buffer / 1e6, ok,
            free_b / 1048576.0);
    fclose(fp);
  }
}

// Status -> the reference's failure behaviour (message, CSV completion, exit).
inline void fail(rmq_status st, const char* what, uint64_t t, uint32_t R, int alg,
                 const CmdArgs& a) {
  if (st == RMQ_ETHRESHOLD) {
    if (alg == RMQ_ALG_VECTOR_LOAD)
      fprintf(stderr, "Error: scan_threshold (%u) must be >= reduction_factor (%u) to avoid edge cases\n",
              (unsigned)t, R);
    else
      fprintf(stderr, "Error: scan_threshold (%u) must be at least 2x reduction_factor (%u) to avoid edge cases\n",
              (unsigned)t, R);
  } else {
    fprintf(stderr, "Error: %s failed: %s\n", what, rmq_status_str(st));
  }
  complete_line(a);
  exit(EXIT_FAILURE);
}

// Queries and GPU BASE by position type.
inline rmq_status query_p(const rmq_hierarchy* h, const uint32_t* d_lr, uint32_t q, float* v, uint32_t* i) {
  return rmq_query(h, d_lr, q, v, i, nullptr);
}
inline rmq_status query_p(const rmq_hierarchy* h, const uint64_t* d_lr, uint32_t q, float* v, uint64_t* i) {
  return rmq_query64(h, d_lr, q, v, i, nullptr);
}
inline rmq_status base_p(const float* d_array, int64_t n, const uint32_t* d_lr, uint32_t q, float* v,
                         uint32_t* i) {
  return rmq_gpu_base(d_array, static_cast<uint64_t>(n), d_lr, q, v, i, nullptr);
}
inline rmq_status base_p(const float* d_array, int64_t n, const uint64_t* d_lr, uint32_t q, float* v,
                         uint64_t* i) {
  return rmq_gpu_base64(d_array, static_cast<uint64_t>(n), d_lr, q, v, i, nullptr);
}

// Shared body of the three algorithm entry points: build, reps x query, report, D2H.
template <typename P>
inline void* run_hierarchical(int alg, int64_t n, uint64_t t, uint32_t q, const float* d_array,
                              const P* d_query, const CmdArgs& a, bool want_idx) {
  rmq_build_options opt{};
  opt.cg_size_log = a.cg_size_log;
  opt.cg_amount_log = a.cg_amount_log;
  opt.group_lanes = a.group_lanes;
  opt.flags = RMQ_BUILD_DEFAULT_CG | (a.pack_descents ? RMQ_BUILD_PACK_DESCENTS : 0u);
  uint32_t R = 0, levels = 0;
  uint64_t buffer_floats = 0;
  rmq_status st = rmq_plan(n, t, alg, &opt, &levels, nullptr, nullptr, &buffer_floats, &R);
  if (st != RMQ_OK) fail(st, "rmq_plan", t, alg == RMQ_ALG_VECTOR_LOAD ? 4u << a.cg_amount_log
                                                  : 1u << (a.cg_size_log + a.cg_amount_log), alg, a);
  void* d_out = nullptr;
  const size_t out_bytes = static_cast<size_t>(q) * (want_idx ? sizeof(P) : 4);
  if (hipMalloc(&d_out, out_bytes ? out_bytes : 4) != hipSuccess) fail(RMQ_ENOMEM, "hipMalloc", t, R, alg, a);

  Timer timer;
  VBHMem mem;
  const unsigned shown_g = alg == RMQ_ALG_VECTOR_LOAD ? 4u : 1u << a.cg_size_log;
  const unsigned shown_c = alg == RMQ_ALG_VECTOR_LOAD ? 1u << a.cg_amount_log : 1u << a.cg_amount_log;
  printf("Build XXX (cg=%2u,amnt=%2u).................", shown_g, shown_c);
  fflush(stdout);
  rmq_hierarchy* h = nullptr;
  timer.restart();
  st = rmq_build_ex(d_array, n, t, alg, &opt, nullptr, &h);
  if (st != RMQ_OK) fail(st, "rmq_build", t, R, alg, a);
  if (hipDeviceSynchronize() != hipSuccess) fail(RMQ_EDEVICE, "build", t, R, alg, a);
  const double build_ms = timer.ms();
  mem.out_buffer = buffer_floats * sizeof(float) * 2;  // values + leftmost indices
  printf("done: %f ms (%u levels, %f MB)\n", build_ms, levels, mem.out_buffer / 1e6);

  printf("Query XXX.................................");
  fflush(stdout);
  timer.restart();
  for (int i = 0; i < a.reps; ++i) {
    st = query_p(h, d_query, q, want_idx ? nullptr : static_cast<float*>(d_out),
                 want_idx ? static_cast<P*>(d_out) : nullptr);
    if (st != RMQ_OK) fail(st, "rmq_query", t, R, alg, a);
    if (hipDeviceSynchronize() != hipSuccess) fail(RMQ_EDEVICE, "query", t, R, alg, a);
  }
  const double time_ms = timer.ms();
  const double avg_s = time_ms / (1000.0 * a.reps);
  printf("done: %f secs (avg %f secs) [%.2f RMQs/sec]\n", time_ms / 1000, avg_s, (double)q / avg_s);
  // COMPLETE_TIME_FILE_CUDA_ERROR, src/device_tools.cuh:5-18
  if (a.save_time && hipGetLastError() != hipSuccess) {
    complete_line(a);
    exit(1);
  }
  write_results(time_ms, q, build_ms, a.reps, a, mem);

  void* out = malloc(out_bytes ? out_bytes : 4);
  if (hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost) != hipSuccess)
    fail(RMQ_EDEVICE, "hipMemcpy", t, R, alg, a);
  (void)hipFree(d_out);
  rmq_free(h);
  return out;
}

template <typename P>
inline float* xxx_rmq(int64_t n, uint64_t t, uint32_t q, const float* d_array, const P* d_query,
                      const CmdArgs& a) {
  return static_cast<float*>(run_hierarchical(RMQ_ALG_XXX, n, t, q, d_array, d_query, a, false));
}

template <typename P>
inline float* xxx_rmq_multi_load(int64_t n, uint64_t t, uint32_t q, const float* d_array,
                                 const P* d_query, const CmdArgs& a) {
  return static_cast<float*>(
      run_hierarchical(RMQ_ALG_XXX_MULTI_LOAD, n, t, q, d_array, d_query, a, false));
}

template <typename P>
inline float* hierarchical_vector_load(int64_t n, uint64_t t, uint32_t q, const float* d_array,
                                       const P* d_query, const CmdArgs& a) {
  return static_cast<float*>(
      run_hierarchical(RMQ_ALG_VECTOR_LOAD, n, t, q, d_array, d_query, a, false));
}

template <typename P>
inline P* hierarchical_rmq_idx(int alg, int64_t n, uint64_t t, uint32_t q, const float* d_array,
                               const P* d_query, const CmdArgs& a) {
  return static_cast<P*>(run_hierarchical(alg, n, t, q, d_array, d_query, a, true));
}

// GPU BASE (algs 2 / 102), gpu_rmq_basic / gpu_rmq_basic_idx, src/cuda_methods.cuh:71-137.
template <typename P>
inline void* gpu_rmq_basic(int64_t n, uint32_t q, const float* d_array, const P* d_query,
                           const CmdArgs& a, bool want_idx) {
  void* d_out = nullptr;
  const size_t out_bytes = static_cast<size_t>(q) * (want_idx ? sizeof(P) : 4);
  if (hipMalloc(&d_out, out_bytes ? out_bytes : 4) != hipSuccess) fail(RMQ_ENOMEM, "hipMalloc", 0, 0, 2, a);
  printf("Computing RMQs (%-16s,r=%-3i)...", want_idx ? "[GPU] BASE IDX" : "[GPU] BASE", a.reps);
  fflush(stdout);
  Timer timer;
  for (int i = 0; i < a.reps; ++i) {
    const rmq_status st = base_p(d_array, n, d_query, q, want_idx ? nullptr : static_cast<float*>(d_out),
                                 want_idx ? static_cast<P*>(d_out) : nullptr);
    if (st != RMQ_OK) fail(st, "rmq_gpu_base", 0, 0, 2, a);
    if (hipDeviceSynchronize() != hipSuccess) fail(RMQ_EDEVICE, "gpu base", 0, 0, 2, a);
  }
  const double ms = timer.ms();
  printf("done: %f secs (avg %f secs): [%.2f RMQs/sec, %f nsec/RMQ]\n", ms / 1000.0,
         ms / (1000.0 * a.reps), (double)q / (ms / 1000.0 / a.reps), ms * 1e6 / a.reps / q);
  write_results(ms, q, 0, a.reps, a, VBHMem{});
  void* out = malloc(out_bytes ? out_bytes : 4);
  (void)hipMemcpy(out, d_out, out_bytes, hipMemcpyDeviceToHost);
  (void)hipFree(d_out);
  return out;
}

// ---- CPU BASE (algs 0 / 100) --------------------------------------------------------------
// Leftmost minimum of A[l..r]: the first element seeds it and a later one replaces it only if
// strictly smaller (cpu_min / cpu_min_idx, src/cpu_methods.h:7-16, 42-53), so the value is
// A[index] bit for bit and a NaN at A[l] is kept.
template <typename P>
inline P cpu_scan_leftmost(const float* A, P l, P r) {
  float m = A[l];
  P at = l;
  for (P i = l + 1; i <= r; ++i) {
    const float x = A[i];
    if (x < m) {
      m = x;
      at = i;
    }
  }
  return at;
}

// reps passes over the q queries (P pairs {l, r}) on the host with a.nt OpenMP threads, timed on
// the host wall clock; returns a malloc'd float[q] of values, or P[q] of indices.
template <typename P>
inline void* cpu_rmq_base(uint32_t q, const float* A, const P* Q, const CmdArgs& a, bool want_idx) {
  omp_set_num_threads(a.nt > 0 ? a.nt : 1);
  if (want_idx)
    printf("Computing RMQs index (%-11s, nt=%2i,r=%-3i)..", "[CPU] BASE IDX", a.nt, a.reps);
  else
    printf("Computing RMQs (%-11s, nt=%2i,r=%-3i)..", "[CPU] BASE", a.nt, a.reps);
  fflush(stdout);
  void* out = malloc(static_cast<size_t>(q) * (want_idx ? sizeof(P) : sizeof(float)) + 8);
  P* oi = static_cast<P*>(out);
  float* ov = static_cast<float*>(out);
  Timer timer;
  for (int rep = 0; rep < a.reps; ++rep) {
#pragma omp parallel for
    for (int64_t k = 0; k < static_cast<int64_t>(q); ++k) {
      const P at = cpu_scan_leftmost<P>(A, Q[2 * k], Q[2 * k + 1]);
      if (want_idx) oi[k] = at;
      else ov[k] = A[at];
    }
  }
  const double ms = timer.ms();
  const double it = ms / a.reps;
  printf("done: %f secs (avg %f secs): [%.2f RMQs/sec, %f nsec/RMQ]\n", ms / 1000.0, ms / (1000.0 * a.reps),
         q / (it / 1000.0), it * 1e6 / q);
  write_results(ms, q, 0, a.reps, a, VBHMem{});
  return out;
}

template <typename P>
inline float* cpu_rmq(uint32_t q, const float* A, const P* Q, const CmdArgs& a) {
  return static_cast<float*>(cpu_rmq_base<P>(q, A, Q, a, false));
}

template <typename P>
inline P* cpu_rmq_idx(uint32_t q, const float* A, const P* Q, const CmdArgs& a) {
  return static_cast<P*>(cpu_rmq_base<P>(q, A, Q, a, true));
}

}  // namespace rtxrmq
