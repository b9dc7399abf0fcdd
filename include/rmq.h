/*
 * rmq.h — C ABI of the MI355X hierarchical range-minimum-query library (librmq.so).
 *
 * This is the drop-in boundary for GPU-RMQ's hierarchical RMQ path (algs 16, 20, 24).
 * In the reference there is no plugin layer: `main` calls a C++ template per algorithm
 * that allocates, builds, queries and frees in one call and returns a host buffer:
 *   xxx_rmq              (src/xxx.h:203-298)   alg 16, "GPU-RMQ (CL)"
 *   hierarchical_vector_load (src/xxx.h:932-1040) alg 20, "GPU-RMQ (VL)"
 *   xxx_rmq_multi_load   (src/xxx.h:530-627)   alg 24, "CL multi load"
 * Here build and query are split so a built hierarchy can stay resident in HBM:
 *   rmq_build  replaces the metadata loop + build launches  (src/xxx.h:211-271)
 *   rmq_query  replaces the query launch                    (src/xxx.h:273-285)
 *   rmq_free   replaces the frees                           (src/xxx.h:294-296)
 *
 * Conventions
 *  - All pointers named d_* are device pointers (hipMalloc / torch tensors).
 *  - `stream` is a hipStream_t passed as void* (NULL = the null stream). Every call is
 *    stream-ordered and asynchronous; nothing on the query path allocates memory.
 *  - Queries are `uint32_t[2*q]` = q pairs {l, r}, inclusive, 0 <= l <= r < n, exactly the
 *    reference's uint2 layout (src/main.cu:25-26, src/rand.cuh:79-80).
 *  - Results are exact: value = A[leftmost argmin], index = leftmost argmin under a strict
 *    `<` scan from l to r, identical to CPU BASE (src/cpu_methods.h:7-16, 42-53). That includes
 *    NaN input: the scan starts from A[l] and a later element replaces the minimum only if
 *    strictly smaller, so a NaN at A[l] answers (A[l], l), and a NaN anywhere else never wins
 *    (a range whose other elements are all NaN answers with A[l]). The build records whether the
 *    array holds a NaN, so queries on NaN-free arrays skip the A[l] test.
 *  - Errors are status codes instead of the reference's exit() (src/xxx.h:211-217).
 *  - A built hierarchy is immutable; concurrent rmq_query calls on different streams are
 *    allowed. The hierarchy borrows d_array: it must stay valid and unchanged until rmq_free.
 */
#ifndef RMQ_AMD_RMQ_H
#define RMQ_AMD_RMQ_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
  RMQ_OK = 0,
  RMQ_EINVAL = -1,        /* bad pointer, size, alignment, alg or option */
  RMQ_ETHRESHOLD = -2,    /* t < 2R (algs 16/24) or t < R (alg 20): src/xxx.h:211-217,938-944 */
  RMQ_ENOMEM = -3,        /* device allocation failed */
  RMQ_EDEVICE = -4,       /* HIP runtime error (launch or API) */
  RMQ_EUNSUPPORTED = -5,  /* option combination not compiled in */
  RMQ_EAGAIN = -6         /* a server's staging ring is full: poll (or drain), then submit again */
} rmq_status;

/* Algorithm ids, the reference's own numbering (src/main.cu:36-53). */
enum {
  RMQ_ALG_GPU_BASE = 2,             /* brute-force scan per query, src/cuda_methods.cuh:4-68 */
  RMQ_ALG_XXX = 16,                 /* GPU-RMQ (CL) */
  RMQ_ALG_VECTOR_LOAD = 20,         /* GPU-RMQ (VL), Hierarchical_vector_load */
  RMQ_ALG_XXX_MULTI_LOAD = 24       /* CL multi load */
};

#define RMQ_MAX_LEVELS 20           /* XXX_MAX_LEVELS, src/constants.cuh:14 */
#define RMQ_LEVEL_ALIGNMENT 32      /* XXX_LEVEL_ALIGNMENT, src/constants.cuh:15 */

typedef struct rmq_hierarchy rmq_hierarchy;

/* Build options. Zero-initialise and set what you need; 0 means "reference default".
 * cg_size_log / cg_amount_log are the reference's compile-time XXX_CG_SIZE_LOG /
 * XXX_CG_AMOUNT_LOG (CMakeLists.txt:61-67, defaults 3 and 2). They fix the reduction
 * factor R of the hierarchy exactly as in the reference:
 *   algs 16/24: R = 2^(cg_size_log + cg_amount_log)      (src/xxx.h:208-209)
 *   alg 20:     R = 2^(cg_amount_log + 2)                 (src/xxx.h:934-935)
 * group_lanes is this library's own execution choice (lanes that cooperate on one
 * query in algs 16/24; 0 = auto; 1 = a lane per query with 16-byte loads, the solver of alg 20,
 * for R <= 64 and not with RMQ_BUILD_FULL; 64 = a whole wavefront per query, the long-scan form
 * for large t). It never changes results. */
typedef struct {
  uint32_t cg_size_log;     /* 0 -> 3 */
  uint32_t cg_amount_log;   /* 0 -> 2 */
  uint32_t group_lanes;     /* 0 -> auto; else 1, 4, 8, 16, 32 or 64 */
  uint32_t flags;           /* RMQ_BUILD_* below */
} rmq_build_options;

#define RMQ_BUILD_DEFAULT_CG 0x1u   /* use cg_size_log = cg_amount_log = 0 literally */
#define RMQ_BUILD_FULL 0x2u         /* read every piece of the traversal in the reference's
                                       order instead of the pruned order (same answers) */
/* A build also keeps a sparse table over the top level (m entries): floor(log2 m) rows of m
 * 8-byte {value, level-0 index} entries, built when m >= 64 and the table is at most 64 MiB, and not
 * with RMQ_BUILD_FULL. A query whose remaining range on the top level is longer than 8 entries
 * reads two table entries instead of scanning it, so large t no longer costs long scans (the
 * reference default t = 2^15 with uniform ranges). The flag below builds none (same answers). */
#define RMQ_BUILD_NO_TOP_TABLE 0x4u
/* Lane-per-query solver (alg 20; algs 16/24 with group_lanes 1) of rmq_query with its descents
 * packed: every lane reads its final range and one level of its tails, and the queries that
 * still descend are regrouped through LDS onto as few wavefronts as they fill, instead of each
 * wavefront walking until its deepest lane is done. Same answers. Measured at n = q = 2^26:
 * +10-12% on lognormal-medium lengths, about even on mixed lengths, 2-6% slower on uniform and
 * short lengths, where the regrouping is not repaid. Not with RMQ_BUILD_FULL; the other solvers
 * and rmq_query64 ignore it. */
#define RMQ_BUILD_PACK_DESCENTS 0x8u
/* Debug validation of queries: rmq_query and rmq_query64 on this hierarchy first check every
 * query (0 <= l <= r < n) on the device and return RMQ_EINVAL, answering nothing, if one is out
 * of range (the reference does not check them, src/main.cu:25-26). The check synchronises the
 * stream and uses one device word of the hierarchy, so checked calls on one hierarchy must not
 * overlap (a call on another stream could see the other call's result). */
#define RMQ_BUILD_CHECK_QUERIES 0x10u
/* Lane-per-query solver with two adjacent queries per lane: both queries' top-level table reads
 * are in flight together (queries answered from the table are a chain of two dependent reads).
 * At R = 4 where the top level holds 16..1024 entries, the kernel keeps the top level and the
 * table rows it reads in LDS (60 KiB per 1024-lane workgroup, two per CU) and parks every query
 * the table does not answer outright in a per-wavefront LDS queue, solved 64 at a time: the
 * fastest form on uniform lengths (257 Gq/s at n = q = 2^26), slower than the packed form below
 * on lognormal and short lengths.
 * With RMQ_BUILD_PACK_DESCENTS as well, both queries read their final ranges, walk one level
 * of their tails where many lanes of the wavefront descend, and the queries still descending
 * are regrouped through LDS (the fastest form on lognormal-medium lengths; uniform lengths as
 * fast as without packing). Same answers. Not with RMQ_BUILD_FULL; the other solvers and
 * rmq_query64 ignore it. */
#define RMQ_BUILD_PAIRED 0x20u
/* Block scans: the build also keeps, for every entry i of the upper levels, the leftmost minimum
 * of the rest of its R-block [i, block end) and of the block up to it [block start, i], each
 * with its level-0 index (16 bytes per entry; levels from the top down while they take at most
 * 64 MiB: all of them at n = 2^26, R = 32, levels 2.. at n = 2^30). Where R >= 16 the pruned
 * solvers (the lane-per-query one: alg 20, algs 16/24 with group_lanes 1; and the G-lane groups
 * of algs 16/24) then read a tail piece above level 0 as one entry, and a final range of more
 * than 16 entries that spans two blocks as two, instead of scanning up to R - 1 (2R - 2)
 * entries. Same answers. Built by default at R >= 32 (the algs 16/24 default);
 * RMQ_BUILD_BLOCK_SCANS builds them at R = 16 too (measured at n = q = 2^26 there for the lane
 * solver: lognormal-medium and uniform lengths faster, short and mixed lengths slower) and
 * RMQ_BUILD_NO_BLOCK_SCANS builds none. Ignored at R <= 8 and by RMQ_BUILD_FULL. */
#define RMQ_BUILD_BLOCK_SCANS 0x40u
#define RMQ_BUILD_NO_BLOCK_SCANS 0x80u
/* Chunk size (log2, 8..30) of the 64-bit form for this build; arrays of more than 2^31 - 1
 * elements use 2^30 without it. Setting it holds even a small array in chunks, which is how the
 * 64-bit form is tested on small inputs. */
#define RMQ_BUILD_CHUNK_LOG(k) ((uint32_t)(k) << 8)

/* Build the hierarchy of block minima over d_array[0..n) on `stream`.
 * Reference counterpart: metadata loop + per-level xxx_build_step launches,
 * src/xxx.h:225-271 (alg 16), 552-599 (alg 24), 959-1010 (alg 20).
 * Requires 1 <= n <= 2^38, d_array 16-byte aligned, t >= 1. Up to 2^31 - 1 (the reference's
 * int32 n) the hierarchy is one piece; above, it is the 64-bit form (the reference's IS_LONG
 * build, src/main.cu:21-27): chunks of 2^30 elements, each its own hierarchy, plus every
 * chunk's minimum, queried with rmq_query64. */
rmq_status rmq_build(const float* d_array, uint64_t n, uint64_t t, int alg,
                     void* stream, rmq_hierarchy** out);

rmq_status rmq_build_ex(const float* d_array, uint64_t n, uint64_t t, int alg,
                        const rmq_build_options* opt, void* stream, rmq_hierarchy** out);

/* Answer q queries. d_val and/or d_idx may be NULL (not both).
 * Reference counterpart: xxx_query (src/xxx.h:124-200), xxx_query_multi_load
 * (src/xxx.h:470-528), kernel_hierarchical_vector_load (src/xxx.h:883-930); the
 * index output has no counterpart there and follows CPU BASE IDX (src/cpu_methods.h:42-53).
 * d_lr must be 8-byte aligned. q = 0 is a no-op. */
rmq_status rmq_query(const rmq_hierarchy* h, const uint32_t* d_lr, uint64_t q,
                     float* d_val, uint32_t* d_idx, void* stream);

/* 64-bit queries: q pairs {l, r} of uint64_t (the reference's IS_LONG ulong2, src/main.cu:21-27,
 * src/tools.h:27-31), 16-byte aligned; indices out as uint64_t. Works on every hierarchy; it is
 * the only query of a 64-bit-form (chunked) one, on which rmq_query returns RMQ_EINVAL. A query
 * across chunks is answered from its first and last chunk's hierarchies and the minima of the
 * chunks in between, with the same leftmost rule. Allocates nothing; concurrent calls on
 * different streams are allowed, as for rmq_query. */
rmq_status rmq_query64(const rmq_hierarchy* h, const uint64_t* d_lr, uint64_t q,
                       float* d_val, uint64_t* d_idx, void* stream);

/* Release the level buffers and the handle. Synchronises the device. NULL is a no-op. */
void rmq_free(rmq_hierarchy* h);

/* Level metadata of a built hierarchy, the reference's xxx_metadata_type
 * (src/constants.cuh:34-39): sizes[k], offsets[k] for k < *num_levels (offsets[0] is the
 * ~0u sentinel of src/xxx.h:229), *buffer_floats = floats in the level buffer, *R = the
 * reduction factor. Any out pointer may be NULL. */
rmq_status rmq_get_metadata(const rmq_hierarchy* h, uint32_t* num_levels, uint64_t* sizes,
                            uint64_t* offsets, uint64_t* buffer_floats, uint32_t* R);

/* Read back the levels of a built hierarchy into caller device buffers of buffer_floats entries,
 * laid out as the reference's level buffer (d_level_buffer, src/xxx.h:244-249): entry j of level
 * k >= 1 at offsets[k] + j. d_val receives the block minima (xxx_build_step, src/xxx.h:91-120);
 * d_idx the leftmost level-0 index of each minimum (no reference counterpart: the reference keeps
 * values only). Either may be NULL, not both. Entries in the alignment gaps between levels are
 * left as they were. Asynchronous on `stream`; RMQ_EUNSUPPORTED for the 64-bit (chunked) form. */
rmq_status rmq_read_levels(const rmq_hierarchy* h, float* d_val, uint32_t* d_idx, void* stream);

/* Same metadata without building anything (host only, no device use). */
rmq_status rmq_plan(uint64_t n, uint64_t t, int alg, const rmq_build_options* opt,
                    uint32_t* num_levels, uint64_t* sizes, uint64_t* offsets,
                    uint64_t* buffer_floats, uint32_t* R);

/* Plain per-array calls for many arrays at once, issued from a C loop: rmq_build_ex for each
 * (d_arrays[i], n[i], t[i]) and rmq_query for each hierarchy. This is the per-array baseline the
 * batched form (rmq_batch.h) is measured against; on error everything built so far is freed. */
rmq_status rmq_build_each(const float* const* d_arrays, const uint64_t* n, const uint64_t* t,
                          int alg, const rmq_build_options* opt, uint32_t count, void* stream,
                          rmq_hierarchy** out);
rmq_status rmq_query_each(rmq_hierarchy* const* hs, const uint32_t* const* d_lr, const uint64_t* q,
                          float* const* d_val, uint32_t* const* d_idx, uint32_t count, void* stream);

/* GPU BASE (algs 2 / 102): one lane per query scans [l, r] serially with strict `<`.
 * Reference counterpart: kernel_rmq_basic / kernel_rmq_basic_idx / value+index variant,
 * src/cuda_methods.cuh:4-68. Used by the rtxrmq driver's --check, as in src/main.cu:194-214. */
rmq_status rmq_gpu_base(const float* d_array, uint64_t n, const uint32_t* d_lr, uint64_t q,
                        float* d_val, uint32_t* d_idx, void* stream);

/* GPU BASE with 64-bit queries and indices (IS_LONG), n <= 2^38. */
rmq_status rmq_gpu_base64(const float* d_array, uint64_t n, const uint64_t* d_lr, uint64_t q,
                          float* d_val, uint64_t* d_idx, void* stream);

/* Human-readable name of a status code. */
const char* rmq_status_str(rmq_status s);

/* Library version string, e.g. "rmq-mi355x 0.1 gfx950". */
const char* rmq_version(void);

#ifdef __cplusplus
}
#endif

#endif /* RMQ_AMD_RMQ_H */
