// rmq_gen.cpp — seeded input generators of the rtxrmq driver (host only).
//
// The reference's inputs are defined by libstdc++'s <random> objects and the order in which
// src/rand.cuh calls them; this file drives the same objects in the same order so that a
// (n, q, lr, seed) tuple yields the same array and queries. Each function names the
// reference lines it follows.

#include "../../include/rmq_gen.h"

#include <cmath>
#include <cstdint>
#include <random>
#include <vector>

namespace {

using u32 = uint32_t;

// Every generator is a template on the position type T: uint32_t for the reference's default
// build, uint64_t for its IS_LONG build (nType, src/main.cu:21-27). The two give the same
// draws while n <= 2^32 (libstdc++ draws both through its 64-bit unsigned type); above, the
// 64-bit distribution combines two 32-bit outputs per draw.

// One (l, r) pair with range length `len` placed uniformly: l ~ U_int[0, n - len],
// r = l + len - 1 (src/rand.cuh:77-80, 98-101, 121-124). A fresh distribution object is
// made per query in the reference; uniform_int_distribution carries no state, so one
// call with the right bounds is the same draw.
template <typename T>
inline void place(std::mt19937& gen, T n, T len, T* out) {
  std::uniform_int_distribution<T> lrand(0, n - 1 - (len - 1));
  const T l = lrand(gen);
  out[0] = l;
  out[1] = l + (len - 1);
}

// Truncate-and-reject lognormal length (src/rand.cuh:118-120, 154-155, 162-163).
template <typename T>
inline T lognormal_len(std::mt19937& gen, std::lognormal_distribution<double>& dist, T n) {
  T len;
  do {
    len = static_cast<T>(dist(gen));
  } while (len <= 0 || len > n);
  return len;
}

// lr > 0 and lr = -1..-5: one std::mt19937 seeded with seed * tid, tid = 0 because the
// reference forces one thread (src/rand.cuh:66-127, 184-185).
template <typename T>
void fill_single_stream(T* out, uint64_t q, int lr, T n, int seed) {
  const int tid = 0;
  std::mt19937 gen(static_cast<std::mt19937::result_type>(seed * tid));
  if (lr > 0) {  // fill_queries_constant, src/rand.cuh:65-84
    for (uint64_t i = 0; i < q; ++i) place<T>(gen, n, static_cast<T>(lr), out + 2 * i);
    return;
  }
  if (lr == -1) {  // fill_queries_uniform, src/rand.cuh:86-105
    std::uniform_int_distribution<T> dist(1, n);
    for (uint64_t i = 0; i < q; ++i) {
      const T len = dist(gen);
      place<T>(gen, n, len, out + 2 * i);
    }
    return;
  }
  // fill_queries_lognormal, src/rand.cuh:107-127, with the scale chosen at
  // src/rand.cuh:199-206 and passed through an `int` parameter.
  int scale = 1;
  if (lr == -2) scale = static_cast<int>(static_cast<T>(std::pow(static_cast<double>(n), 0.6)));
  if (lr == -3) scale = static_cast<int>(static_cast<T>(std::pow(static_cast<double>(n), 0.3)));
  if (lr == -4) scale = static_cast<int>(std::max<T>(1, n / (1u << 8)));
  if (lr == -5) scale = static_cast<int>(std::max<T>(1, n / (1u << 15)));
  std::lognormal_distribution<double> dist(std::log(scale), 0.3);
  for (uint64_t i = 0; i < q; ++i) {
    const T len = lognormal_len<T>(gen, dist, n);
    place<T>(gen, n, len, out + 2 * i);
  }
}

// lr = -6: fill_queries_mixed, src/rand.cuh:129-170. One std::mt19937(seed); per query a
// category U_int{0,1,2} picks uniform / medium lognormal / small lognormal.
template <typename T>
void fill_mixed(T* out, uint64_t q, T n, int seed) {
  std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
  std::uniform_int_distribution<int> category(0, 2);
  std::uniform_int_distribution<T> large(1, n);
  std::lognormal_distribution<double> medium(
      std::log(static_cast<T>(std::pow(static_cast<double>(n), 0.6))), 0.3);
  std::lognormal_distribution<double> small(
      std::log(static_cast<T>(std::pow(static_cast<double>(n), 0.3))), 0.3);
  for (uint64_t i = 0; i < q; ++i) {
    const int c = category(gen);
    T len;
    if (c == 0) {
      len = large(gen);
    } else if (c == 1) {
      len = lognormal_len<T>(gen, medium, n);
    } else {
      len = lognormal_len<T>(gen, small, n);
    }
    place<T>(gen, n, len, out + 2 * i);
  }
}

constexpr uint64_t kTrivialDistinct = 1u << 10;  // TRIVCHECK_LOG_NUM_DIFFQ, src/main.cu:30

template <typename T>
int gen_queries(uint64_t q, int32_t lr, T n, int32_t seed, int trivial_check, T* out) {
  if (lr == 0 || lr < -6 || (lr > 0 && static_cast<uint64_t>(lr) > n)) return -1;
  // src/rand.cuh:187-192: with trivialCheck only 2^10 distinct queries are drawn.
  const bool tile = trivial_check && q > kTrivialDistinct;
  const uint64_t qgen = tile ? kTrivialDistinct : q;
  if (lr == -6) {
    fill_mixed<T>(out, qgen, n, seed);
  } else {
    fill_single_stream<T>(out, qgen, lr, n, seed);
  }
  if (tile) {  // fill_repeatedly, src/rand.cuh:172-179
    for (uint64_t i = qgen; i < q; ++i) {
      const uint64_t src = i & (kTrivialDistinct - 1);
      out[2 * i] = out[2 * src];
      out[2 * i + 1] = out[2 * src + 1];
    }
  }
  return 0;
}

}  // namespace

extern "C" void rmq_gen_array(uint64_t n, int32_t seed, float* out) {
  // random_array_par_cpu, src/rand.cuh:49-61
  std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
  std::uniform_real_distribution<float> dist(0, 1);
  for (uint64_t i = 0; i < n; ++i) out[i] = dist(gen);
}

extern "C" int rmq_gen_queries(uint64_t q, int32_t lr, uint64_t n, int32_t seed,
                               int trivial_check, uint32_t* out) {
  if (n == 0 || n > 0xFFFFFFFFull) return -1;
  return gen_queries<u32>(q, lr, static_cast<u32>(n), seed, trivial_check, out);
}

extern "C" int rmq_gen_queries64(uint64_t q, int32_t lr, uint64_t n, int32_t seed,
                                 int trivial_check, uint64_t* out) {
  if (n == 0 || n > (1ull << 40)) return -1;
  return gen_queries<uint64_t>(q, lr, n, seed, trivial_check, out);
}

extern "C" uint64_t rmq_gen_sample_indices(uint64_t original_size, uint64_t sample_size,
                                           int32_t seed, uint32_t* out) {
  // randomly_sample_indices, src/rand.cuh:219-240
  if (original_size <= sample_size) {
    for (uint64_t i = 0; i < original_size; ++i) out[i] = static_cast<u32>(i);
    return original_size;
  }
  std::mt19937 gen(static_cast<std::mt19937::result_type>(seed));
  std::uniform_int_distribution<u32> dis(0, static_cast<u32>(original_size - 1));
  for (uint64_t i = 0; i < sample_size; ++i) out[i] = dis(gen);
  return sample_size;
}
