/*
 * rmq_gen.h — host-side input generators of the rtxrmq driver (part of librmq.so, no GPU use).
 *
 * They reproduce the reference's seeded inputs bit-for-bit by driving the same libstdc++
 * <random> objects in the same order:
 *   rmq_gen_array           <- random_array_par_cpu   (src/rand.cuh:49-61)
 *   rmq_gen_queries         <- random_queries_par_cpu (src/rand.cuh:181-217) and the
 *                              fill_queries_* family   (src/rand.cuh:65-170)
 *   rmq_gen_sample_indices  <- randomly_sample_indices (src/rand.cuh:219-240)
 * Reference quirks kept on purpose (they define "the same seeded inputs"):
 *   - queries for lr > 0 and lr = -1..-5 use std::mt19937(seed * tid) with the thread
 *     count forced to 1, so the stream is std::mt19937(0) whatever the seed is
 *     (src/rand.cuh:70,91,112,184-185); only lr = -6 uses std::mt19937(seed) (src/rand.cuh:131);
 *   - lognormal lengths are truncated and rejected outside [1, n] (src/rand.cuh:118-120);
 *   - the lognormal scales are (uint32)pow(n, 0.6) and (uint32)pow(n, 0.3) evaluated by libm
 *     (src/rand.cuh:200,202,135-136);
 *   - trivial_check tiles 2^10 distinct queries (src/rand.cuh:187-215).
 */
#ifndef RMQ_AMD_RMQ_GEN_H
#define RMQ_AMD_RMQ_GEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Fill out[0..n) with std::uniform_real_distribution<float>(0,1) over std::mt19937(seed). */
void rmq_gen_array(uint64_t n, int32_t seed, float* out);

/* Fill out[0..2q) with q inclusive {l, r} pairs for range-length mode lr
 * (lr > 0 constant length; -1 uniform; -2 lognormal medium; -3 lognormal small;
 *  -4, -5 lognormal with scale n/2^8, n/2^15; -6 mixed of -1/-2/-3).
 * trivial_check != 0 generates 2^10 distinct queries and tiles them.
 * Returns 0, or -1 for an unknown lr or lr > n (src/tools.h:115-118). */
int rmq_gen_queries(uint64_t q, int32_t lr, uint64_t n, int32_t seed, int trivial_check,
                    uint32_t* out);

/* The same with 64-bit positions, as the reference's IS_LONG build draws them (nType =
 * uint64_t, src/main.cu:21-27): out[0..2q) as uint64_t. Equal to rmq_gen_queries while
 * n <= 2^32; n up to 2^40. */
int rmq_gen_queries64(uint64_t q, int32_t lr, uint64_t n, int32_t seed, int trivial_check,
                      uint64_t* out);

/* Draw sample_size indices in [0, original_size) (or the identity if sample_size >=
 * original_size); returns the number written. */
uint64_t rmq_gen_sample_indices(uint64_t original_size, uint64_t sample_size, int32_t seed,
                                uint32_t* out);

#ifdef __cplusplus
}
#endif

#endif /* RMQ_AMD_RMQ_GEN_H */
