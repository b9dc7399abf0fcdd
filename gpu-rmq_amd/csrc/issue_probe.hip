// issue_probe.hip — measures the VALU and SALU issue rate of the device (wave-instructions per
// second), the peak that bench.py's instruction-issue roofline divides by.
//
//   issue_probe            prints one JSON line per instruction kind and a summary line
//
// Each kernel runs eight independent dependency chains of one instruction kind per lane (so no
// instruction waits on its predecessor's result), 128 instructions per loop iteration, on 8
// wavefronts per SIMD of every CU; the loop's own SALU compare and branch are the only other
// instructions. The rate is wave-instructions divided by the kernel time of HIP events over
// several launches; every lane stores one word at the end so that nothing is optimised away.
// The instruction counts are those of the source (SQ_INSTS_VALU of the same launch agrees).

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>

namespace {

constexpr int kBlock = 256;
constexpr int kChains = 8;
constexpr int kPerIter = 128;  // instructions per loop iteration (kChains x 16)

enum Kind { kAddU32 = 0, kMinF32 = 1, kCndmask = 2, kSaluAdd = 3 };

__global__ __launch_bounds__(kBlock) void k_issue(uint32_t* __restrict__ out, uint32_t iters, uint32_t seed,
                                                  int kind) {
  uint32_t a[kChains];
  const uint32_t b = seed ^ threadIdx.x;
#pragma unroll
  for (int j = 0; j < kChains; ++j) a[j] = threadIdx.x * (j + 3) + seed;
  if (kind == kAddU32) {
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) asm volatile("v_add_u32 %0, %0, %1" : "+v"(a[j]) : "v"(b));
      }
    }
  } else if (kind == kMinF32) {
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) asm volatile("v_min_f32 %0, %0, %1" : "+v"(a[j]) : "v"(b));
      }
    }
  } else if (kind == kCndmask) {
    const uint64_t mask = __ballot(a[0] < b);  // a lane mask in an SGPR pair
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) asm volatile("v_cndmask_b32 %0, %0, %1, %2" : "+v"(a[j]) : "v"(b), "s"(mask));
      }
    }
  } else {
    uint32_t s[kChains];
#pragma unroll
    for (int j = 0; j < kChains; ++j) s[j] = __builtin_amdgcn_readfirstlane(a[j]);
    const uint32_t sb = __builtin_amdgcn_readfirstlane(b);
    for (uint32_t i = 0; i < iters; ++i) {
#pragma unroll
      for (int u = 0; u < 16; ++u) {
#pragma unroll
        for (int j = 0; j < kChains; ++j) asm volatile("s_mov_b32 %0, %1" : "+s"(s[j]) : "s"(sb ^ j));
      }
    }
#pragma unroll
    for (int j = 0; j < kChains; ++j) a[j] += s[j];
  }
  uint32_t x = 0;
#pragma unroll
  for (int j = 0; j < kChains; ++j) x ^= a[j];
  out[blockIdx.x * kBlock + threadIdx.x] = x;
}

}  // namespace

int main() {
  int dev = 0, cus = 0, clock_khz = 0;
  if (hipGetDevice(&dev) != hipSuccess || hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) {
    fprintf(stderr, "issue_probe: no HIP device\n");
    return 1;
  }
  (void)hipDeviceGetAttribute(&clock_khz, hipDeviceAttributeClockRate, dev);
  const uint32_t blocks = static_cast<uint32_t>(cus) * 8;  // 8 waves per SIMD: 4 SIMDs x 8 waves / 4 waves per block
  uint32_t* out = nullptr;
  if (hipMalloc(reinterpret_cast<void**>(&out), static_cast<size_t>(blocks) * kBlock * 4) != hipSuccess) return 1;
  const char* names[] = {"v_add_u32", "v_min_f32", "v_cndmask_b32", "s_mov_b32"};
  const uint32_t iters = 4096;
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  double best_valu = 0, best_salu = 0;
  for (int kind = 0; kind < 4; ++kind) {
    k_issue<<<blocks, kBlock>>>(out, iters, 12345u, kind);  // warm-up
    (void)hipDeviceSynchronize();
    const int reps = 5;
    (void)hipEventRecord(e0, nullptr);
    for (int r = 0; r < reps; ++r) k_issue<<<blocks, kBlock>>>(out, iters, 12345u + r, kind);
    (void)hipEventRecord(e1, nullptr);
    if (hipEventSynchronize(e1) != hipSuccess) return 1;
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    const double waves = static_cast<double>(blocks) * (kBlock / 64);
    const double insts = waves * iters * kPerIter * reps;
    const double rate = insts / (ms * 1e-3) / 1e9;
    const double per_cu_cycle = clock_khz ? rate * 1e9 / (cus * (clock_khz * 1e3)) : 0.0;
    printf("{\"kind\": \"%s\", \"ms_per_launch\": %.4f, \"ginst_per_s\": %.1f, \"per_cu_per_cycle\": %.3f}\n",
           names[kind], ms / reps, rate, per_cu_cycle);
    if (kind < 3 && rate > best_valu) best_valu = rate;
    if (kind == 3) best_salu = rate;
  }
  printf("{\"cus\": %d, \"clock_mhz\": %.0f, \"valu_peak_ginst_per_s\": %.1f, \"salu_peak_ginst_per_s\": %.1f}\n", cus,
         clock_khz / 1e3, best_valu, best_salu);
  (void)hipFree(out);
  return 0;
}
