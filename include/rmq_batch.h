/*
 * rmq_batch.h — batched form of the hierarchical RMQ (part of librmq.so).
 *
 * No reference counterpart: the reference answers one array per call (src/xxx.h:203-298) and
 * rebuilds per call. Here thousands of independent arrays of uneven length are built with one
 * kernel launch per hierarchy level (over all arrays at once) and their uneven query batches
 * are answered by ONE persistent kernel launch that pulls 256-query tiles from a device-side
 * work queue. Every array has its own scan threshold t and execution variant, chosen by the
 * caller (e.g. from the persisted autotuner table); the variant fixes R: 4 for RMQ_VARIANT_LANE4
 * (the reference's VL with cg_amount_log 0), 8 for RMQ_VARIANT_LANE8 (the reference's VL paper
 * setting), 32 (the CL default) otherwise. Results
 * are identical to rmq_build + rmq_query per array (value = A[leftmost argmin], leftmost index).
 *
 * Layout: all arrays live in one device buffer d_arrays; array a occupies
 * d_arrays[offset_a .. offset_a + n_a) with offset_a a multiple of 4. Queries of all arrays are
 * concatenated: array a's q_a pairs start at pair index qoffset_a and hold LOCAL (l, r) pairs
 * (0 <= l <= r < n_a); outputs use the same global query index.
 */
#ifndef RMQ_AMD_RMQ_BATCH_H
#define RMQ_AMD_RMQ_BATCH_H

#include <stdint.h>

#include "rmq.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Execution variants (per array; results never depend on them). */
enum {
  RMQ_VARIANT_GROUP4 = 4,    /* 4 lanes cooperate on one query */
  RMQ_VARIANT_GROUP8 = 8,    /* 8 lanes: one 16-byte vector per lane per 32-float block */
  RMQ_VARIANT_GROUP16 = 16,  /* 16 lanes */
  RMQ_VARIANT_GROUP32 = 32,  /* 32 lanes */
  RMQ_VARIANT_WAVE = 64,     /* a whole wavefront per query (long final ranges, large t) */
  RMQ_VARIANT_LANE = 1,      /* one lane per query, 16-byte loads, R = 32 hierarchy */
  RMQ_VARIANT_LANE8 = 2,     /* one lane per query, 16-byte loads, R = 8 hierarchy */
  RMQ_VARIANT_LANE4 = 3      /* one lane per query, 16-byte loads, R = 4 hierarchy */
};

typedef struct {
  uint64_t offset;   /* element offset of the array in d_arrays (multiple of 4) */
  uint32_t n;        /* 1 <= n <= 2^31 - 1 */
  uint32_t t;        /* scan threshold: >= 64 (2R at R = 32); >= 8 for RMQ_VARIANT_LANE8, >= 4 for
                        RMQ_VARIANT_LANE4 (R) */
  uint32_t variant;  /* RMQ_VARIANT_* */
  uint32_t reserved;
} rmq_batch_array;

typedef struct rmq_batch rmq_batch;
typedef struct rmq_batch_plan rmq_batch_plan;

/* Build every array's hierarchy: one launch per level over all arrays. d_arrays is borrowed
 * until rmq_batch_free. */
rmq_status rmq_batch_build(const float* d_arrays, const rmq_batch_array* arrays, uint32_t count,
                           void* stream, rmq_batch** out);

/* Upload the query tiling for a query layout: qoffsets[count + 1] (host) gives each array's
 * first pair index in the concatenated query buffer (qoffsets[count] = total pairs). The plan
 * owns the device work queue (tiles + counter); it is reusable for any query contents with
 * the same layout. */
rmq_status rmq_batch_plan_queries(const rmq_batch* b, const uint64_t* qoffsets, void* stream,
                                  rmq_batch_plan** out);

/* Answer all queries with one persistent launch (plus a one-lane counter reset launch, so the
 * pair can be captured in a HIP graph). d_val and/or d_idx may be NULL (not both). */
rmq_status rmq_batch_query(const rmq_batch* b, const rmq_batch_plan* p, const uint32_t* d_lr,
                           float* d_val, uint32_t* d_idx, void* stream);

/* rmq_batch_query without a plan, for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). No counterpart in the
 * reference. The query layout comes as d_qoffsets, a DEVICE array of count + 1 entries read like
 * rmq_batch_plan_queries's qoffsets (array a's pairs start at pair d_qoffsets[a]; d_qoffsets[count]
 * ends the last), and total_queries = d_qoffsets[count] - d_qoffsets[0]. The work queue (tiles,
 * queue starts, counters) is built on the device, by two small kernels before the persistent
 * query kernel, in a static array of the library: about 4.3 MiB per GPU, held until the process
 * ends. Nothing is allocated and nothing is copied; every step is a kernel launch on `stream`, so
 * the call can be captured in a HIP graph. Answers equal rmq_batch_query's.
 * Limits: count <= RMQ_BATCH_NOALLOC_MAX_ARRAYS and total_queries / 256 + count <=
 * RMQ_BATCH_NOALLOC_MAX_TILES (2^26 queries in 2^16 arrays, for example); larger calls return
 * RMQ_EINVAL before anything runs. Offsets that do not add up to total_queries answer nothing.
 * Every call on one GPU shares that GPU's static array and it keeps the last call's contents:
 * calls that overlap on one GPU (other streams or threads) race and must be serialised by the
 * caller. */
#define RMQ_BATCH_NOALLOC_MAX_ARRAYS 65536u
#define RMQ_BATCH_NOALLOC_MAX_TILES 262144u
rmq_status rmq_batch_query_noalloc(const rmq_batch* b, const uint64_t* d_qoffsets, uint64_t total_queries,
                                   const uint32_t* d_lr, float* d_val, uint32_t* d_idx, void* stream);

/* Capture rmq_batch_query into a HIP graph on `stream` and instantiate it; replay with
 * rmq_graph_launch. Buffers are bound at capture time. */
typedef struct rmq_graph rmq_graph;
rmq_status rmq_batch_capture(const rmq_batch* b, const rmq_batch_plan* p, const uint32_t* d_lr,
                             float* d_val, uint32_t* d_idx, void* stream, rmq_graph** out);
rmq_status rmq_graph_launch(const rmq_graph* g, void* stream);
void rmq_graph_free(rmq_graph* g);

void rmq_batch_plan_free(rmq_batch_plan* p);
void rmq_batch_free(rmq_batch* b);

/* Bytes of device memory the batch's level buffers occupy. */
uint64_t rmq_batch_level_bytes(const rmq_batch* b);

/* Read back array `array`'s levels as rmq_read_levels does for a plain hierarchy: entry j of level
 * k >= 1 at offsets[k] + j of that array's own metadata (rmq_plan with its variant's alg and
 * options), into device buffers of its buffer_floats entries. Reference counterpart: the level
 * buffer xxx_build_step writes (src/xxx.h:91-120, 244-271), here one array's part of the batch. */
rmq_status rmq_batch_read_levels(const rmq_batch* b, uint32_t array, float* d_val, uint32_t* d_idx,
                                 void* stream);

/* Fill d_out[0..n) with uniform floats in [0, 1): SplitMix64 of seed + (start + i) * golden
 * gamma, top 24 bits scaled by 2^-24 (exact in float32). Used for the synthetic batched and
 * serving workloads. */
rmq_status rmq_fill_uniform(float* d_out, uint64_t n, uint64_t seed, uint64_t start, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* RMQ_AMD_RMQ_BATCH_H */
