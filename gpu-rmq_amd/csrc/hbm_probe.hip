// hbm_probe.hip — measures what this device's HBM delivers to plain kernels, the achievable
// ceilings bench.py reports the query and build kernels against (beside the 8 TB/s of the
// data sheet):
//
//   read       grid-stride 16-byte loads over 4 GiB, summed (one word stored per lane)
//   read_nt    the same with non-temporal loads
//   copy       16-byte loads stored to a second 4 GiB buffer (bytes read + written)
//   read4w1    16-byte loads, one float (their minimum) stored per vector: the level-0 build mix
//   random128  independent random 128-byte segments, eight lanes per segment (segments per
//              second: the rate of random HBM line fetches, what a latency-bound query kernel
//              issues)
//
//   hbm_probe   prints one JSON line per kernel and grid, then a summary line of the best
//
// Every index is below the buffer's element count: the grid-stride loops compare with it, and
// random segments are reduced modulo the segment count. Times are HIP events around each launch,
// best of four after one warm-up.

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>

namespace {

constexpr int kBlock = 256;
constexpr uint64_t kBytes = 4ull << 30;
constexpr uint64_t kVec = kBytes / 16;  // float4 elements per buffer
typedef float f32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(kBlock) void k_read(const float4* __restrict__ a, uint64_t n, float* __restrict__ out) {
  float s = 0.f;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const float4 v = a[i];
    s += v.x + v.y + v.z + v.w;
  }
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_read_nt(const float4* __restrict__ a, uint64_t n, float* __restrict__ out) {
  float s = 0.f;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const f32x4 v = __builtin_nontemporal_load(reinterpret_cast<const f32x4*>(a) + i);
    s += v.x + v.y + v.z + v.w;
  }
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}

__global__ __launch_bounds__(kBlock) void k_copy(const float4* __restrict__ a, uint64_t n, float4* __restrict__ b) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) b[i] = a[i];
}

__global__ __launch_bounds__(kBlock) void k_read4w1(const float4* __restrict__ a, uint64_t n, float* __restrict__ b) {
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < n; i += uint64_t(gridDim.x) * kBlock) {
    const float4 v = a[i];
    b[i] = fminf(fminf(v.x, v.y), fminf(v.z, v.w));
  }
}

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// segs random segments of 8 float4 (128 bytes), eight consecutive lanes per segment
__global__ __launch_bounds__(kBlock) void k_random128(const float4* __restrict__ a, uint64_t nseg, uint64_t segs,
                                                      uint64_t salt, float* __restrict__ out) {
  float s = 0.f;
  for (uint64_t i = blockIdx.x * uint64_t(kBlock) + threadIdx.x; i < segs * 8; i += uint64_t(gridDim.x) * kBlock) {
    const uint64_t seg = mix64((i >> 3) + salt) % nseg;
    const float4 v = a[seg * 8 + (i & 7)];
    s += v.x + v.w;
  }
  out[blockIdx.x * kBlock + threadIdx.x] = s;
}

bool ok(hipError_t e, const char* what) {
  if (e == hipSuccess) return true;
  std::fprintf(stderr, "hbm_probe: %s: %s\n", what, hipGetErrorString(e));
  return false;
}

}  // namespace

int main() {
  constexpr int kGrids[3] = {4096, 16384, 65536};
  constexpr uint64_t kSegs = 1ull << 26;
  float4 *a = nullptr, *b = nullptr;
  float* sink = nullptr;
  if (!ok(hipMalloc(&a, kBytes), "hipMalloc") || !ok(hipMalloc(&b, kBytes), "hipMalloc") ||
      !ok(hipMalloc(&sink, sizeof(float) * 65536 * kBlock), "hipMalloc") ||
      !ok(hipMemset(a, 0, kBytes), "hipMemset") || !ok(hipMemset(b, 0, kBytes), "hipMemset"))
    return 1;
  hipEvent_t e0, e1;
  if (!ok(hipEventCreate(&e0), "hipEventCreate") || !ok(hipEventCreate(&e1), "hipEventCreate")) return 1;
  const char* names[5] = {"read", "read_nt", "copy", "read4w1", "random128"};
  double best[5] = {0, 0, 0, 0, 0};  // GB/s, or G segments/s for random128
  for (int k = 0; k < 5; ++k) {
    for (int grid : kGrids) {
      float ms_best = 1e30f;
      for (int r = 0; r < 5; ++r) {
        if (!ok(hipEventRecord(e0), "hipEventRecord")) return 1;
        switch (k) {
          case 0: k_read<<<grid, kBlock>>>(a, kVec, sink); break;
          case 1: k_read_nt<<<grid, kBlock>>>(a, kVec, sink); break;
          case 2: k_copy<<<grid, kBlock>>>(a, kVec, b); break;
          case 3: k_read4w1<<<grid, kBlock>>>(a, kVec, reinterpret_cast<float*>(b)); break;
          default: k_random128<<<grid, kBlock>>>(a, kVec / 8, kSegs, static_cast<uint64_t>(r) << 40, sink); break;
        }
        if (!ok(hipGetLastError(), "launch") || !ok(hipEventRecord(e1), "hipEventRecord") ||
            !ok(hipEventSynchronize(e1), "hipEventSynchronize"))
          return 1;
        float ms = 0.f;
        if (!ok(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime")) return 1;
        if (r > 0 && ms < ms_best) ms_best = ms;
      }
      const double s = ms_best / 1e3;
      double rate;
      if (k == 4) rate = kSegs / s / 1e9;
      else rate = (k == 2 ? 2.0 : (k == 3 ? 1.25 : 1.0)) * kBytes / s / 1e9;
      if (rate > best[k]) best[k] = rate;
      std::printf("{\"kernel\": \"%s\", \"grid\": %d, \"ms\": %.4f, \"%s\": %.2f}\n", names[k], grid, ms_best,
                  k == 4 ? "gsegments_per_s" : "gbps", rate);
      std::fflush(stdout);
    }
  }
  std::printf("{\"read_gbps\": %.1f, \"read_nt_gbps\": %.1f, \"copy_gbps\": %.1f, \"read4w1_gbps\": %.1f, "
              "\"random128_gsegments_per_s\": %.2f}\n",
              best[0], best[1], best[2], best[3], best[4]);
  (void)hipFree(a);
  (void)hipFree(b);
  (void)hipFree(sink);
  return 0;
}
