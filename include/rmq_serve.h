/*
 * rmq_serve.h — serving loop over device-resident hierarchies (part of librmq.so).
 *
 * No reference counterpart (the reference answers one batch per process run, src/main.cu).
 * A server keeps built hierarchies resident in HBM, admits new arrays only while they fit in
 * free device memory (minus a reserve), and answers requests by continuous batching: pending
 * requests (each: queries on one resident array) are packed into the next free "slot" — a
 * stream with its own tile buffers — and answered by one persistent multi-array launch
 * (the batched form's kernel, rmq_batch.h); up to `slots` batches are in flight at once, so
 * requests that arrive while others are on the GPU overlap with them. Answers are identical to
 * rmq_query on the same hierarchy, whatever else is in the batch.
 *
 * Staging: submit copies (and checks) a request's pairs once, into a ring of pinned host memory
 * that the kernels read in place; they write the answers into a ring beside it, from which poll
 * copies them to the caller's buffers. A request holds its ring space until it is reported done;
 * when the ring has no room, submit returns RMQ_EAGAIN.
 *
 * The server is driven by the caller (rmq_server_poll / rmq_server_drain) from one thread; it
 * starts no threads of its own.
 */
#ifndef RMQ_AMD_RMQ_SERVE_H
#define RMQ_AMD_RMQ_SERVE_H

#include <stdint.h>

#include "rmq.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct rmq_server rmq_server;

typedef struct {
  uint32_t slots;              /* batches in flight at once; 0 -> 3 */
  uint32_t max_batch_queries;  /* queries one batch holds; 0 -> 2^20 (a request may not exceed it) */
  uint64_t reserve_bytes;      /* free HBM admission keeps untouched; 0 -> 2 GiB */
  uint32_t max_arrays;         /* resident arrays; 0 -> 1024 */
  uint32_t max_batch_requests; /* requests one batch holds; 0 -> 4096 */
  uint64_t ring_queries;       /* staging ring, in queries (12 bytes of pinned host memory each:
                                  the pair and the answer); 0 -> max(2^22, 2 x slots x
                                  max_batch_queries); at least max_batch_queries */
} rmq_server_options;

enum {
  RMQ_ADMIT_DEVICE = 0,   /* `data` is a device array, borrowed until eviction; admit first waits
                             for all work on the device, so writes to it issued before the call
                             on any stream are complete when it is built */
  RMQ_ADMIT_HOST = 1,     /* `data` is a host array, copied into server-owned device memory */
  RMQ_ADMIT_UNIFORM = 2   /* server-owned device memory filled by rmq_fill_uniform(seed) */
};

typedef struct {
  uint64_t submitted, completed, batches, queries;
  uint64_t resident_bytes;     /* device bytes owned by resident arrays and their levels */
  uint32_t resident, in_flight;
} rmq_server_stats;

rmq_status rmq_server_create(const rmq_server_options* opt, rmq_server** out);

/* Waits for in-flight batches, frees everything the server owns. NULL is a no-op. */
void rmq_server_destroy(rmq_server* s);

/* Build and keep resident an array of n floats with scan threshold t and execution variant
 * RMQ_VARIANT_* (R = 4 for RMQ_VARIANT_LANE4, t >= 4; R = 8 for RMQ_VARIANT_LANE8, t >= 8;
 * R = 32 otherwise, t >= 64). Returns RMQ_ENOMEM, with nothing allocated, when the array (if the
 * server owns it) and its levels do not fit in free HBM minus the reserve. */
rmq_status rmq_server_admit(rmq_server* s, int source, const float* data, uint64_t n, uint64_t seed,
                            uint32_t t, uint32_t variant, int32_t* array_id);

/* Drop a resident array. RMQ_EINVAL if it is unknown or a pending/in-flight request uses it. */
rmq_status rmq_server_evict(rmq_server* s, int32_t array_id);

/* Queue q queries (host {l, r} pairs, copied now into the staging ring) on a resident array.
 * Answers are written to the caller's host buffers h_val / h_idx (either may be NULL, not both)
 * before the ticket is reported done. q must be 1..max_batch_queries. Every pair must satisfy
 * 0 <= l <= r < n of the array: the pairs are checked while they are copied, and a request
 * holding any other pair is refused with RMQ_EINVAL, with nothing queued (the batch kernel is
 * shared by all requests in flight, so a bad pair is never handed to it). RMQ_EAGAIN, with
 * nothing queued, when the ring has no room for q pairs until earlier requests are done. */
rmq_status rmq_server_submit(rmq_server* s, int32_t array_id, const uint32_t* h_lr, uint64_t q,
                             float* h_val, uint32_t* h_idx, uint64_t* ticket);

/* Non-blocking step: retire finished batches (copy answers out), then pack pending requests
 * into free slots and launch them. Up to max_done finished tickets are written to done. */
rmq_status rmq_server_poll(rmq_server* s, uint64_t* done, uint32_t max_done, uint32_t* n_done);

/* Block until no request is pending or in flight (tickets as in rmq_server_poll). */
rmq_status rmq_server_drain(rmq_server* s, uint64_t* done, uint32_t max_done, uint32_t* n_done);

rmq_status rmq_server_get_stats(const rmq_server* s, rmq_server_stats* st);

#ifdef __cplusplus
}
#endif

#endif /* RMQ_AMD_RMQ_SERVE_H */
