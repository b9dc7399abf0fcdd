// oracle.cpp — CPU restatement of GPU-RMQ's hierarchical RMQ path. TEST INFRASTRUCTURE ONLY.
//
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load this library,
// and only as the checker (or the timed CPU baseline). The product (librmq.so, rtxrmq) never
// links or calls it.
//
// Functions and the reference lines they restate:
//   oracle_cpu_base_value  CPU BASE, alg 0      src/cpu_methods.h:7-16 (cpu_min), 18-39 (cpu_rmq)
//   oracle_cpu_base_index  CPU BASE IDX, alg 100 src/cpu_methods.h:42-53 (cpu_min_idx), 55-77
//   oracle_plan            level metadata        src/xxx.h:225-243 (also 393-411, 552-570, 959-977)
//   oracle_build           level contents        src/xxx.h:91-120 (xxx_build_step), plus the
//                                                leftmost argmin of every block (no reference
//                                                counterpart: the reference keeps values only)
//   oracle_hier_query      hierarchical query    src/xxx.h:124-200 (alg 16), 470-528 (alg 24),
//                                                883-930 (alg 20)
//   oracle_cover           traversal coverage    the same traversals, recording which level-0
//                                                elements each visited piece stands for
//
// How each is pinned (tests/test_oracle.py): CPU BASE against numpy min/argmin (numpy's argmin
// returns the first occurrence, i.e. the leftmost rule) and hand-made tie/±0/inf vectors;
// metadata against the closed-form known answers of the recurrence; level contents against a
// numpy block-min; the traversals by exact coverage of [l, r] plus brute-force equality.
//
// Build: oracle/Makefile (g++ -O3 -fopenmp -shared).

#include <omp.h>

#include <cmath>
#include <cstdint>
#include <cstring>

// ORACLE_MUTANT plants one known mistake; only tests/test_oracle.py builds such variants, to
// show that the checks pinning the oracle fail on them. 0 (the default) is the oracle proper.
#ifndef ORACLE_MUTANT
#define ORACLE_MUTANT 0
#endif

namespace {

using u32 = uint32_t;
using u64 = uint64_t;

constexpr u32 kMaxLevels = 20;      // XXX_MAX_LEVELS, src/constants.cuh:14
constexpr u64 kAlignment = 32;      // XXX_LEVEL_ALIGNMENT, src/constants.cuh:15
constexpr u32 kNoIdx = 0xFFFFFFFFu;

// Strict `<` scan from l to r (inclusive), src/cpu_methods.h:8-15 and 43-52.
inline u32 cpu_min_idx(const float* A, u32 l, u32 r) {
#if ORACLE_MUTANT == 7
  float m = INFINITY;  // planted: the scan is seeded with +inf instead of A[l]
#else
  float m = A[l];
#endif
  u32 mi = l;
  for (u64 i = l; i <= r; ++i) {
#if ORACLE_MUTANT == 1
    if (A[i] <= m) {  // planted: rightmost tie
#else
    if (A[i] < m) {
#endif
      m = A[i];
      mi = static_cast<u32>(i);
    }
  }
  return mi;
}

inline float cpu_min(const float* A, u32 l, u32 r) {
#if ORACLE_MUTANT == 7
  float m = INFINITY;
#else
  float m = A[l];
#endif
#if ORACLE_MUTANT == 2
  for (u64 i = l; i < r; ++i) {  // planted: drops the last element
#else
  for (u64 i = l; i <= r; ++i) {
#endif
    if (A[i] < m) m = A[i];
  }
  return m;
}

// Candidate ordering used by the hierarchy: smaller value first, then smaller index.
// With strict `<` scanning left to right, CPU BASE picks exactly this candidate.
struct Best {
  float v = INFINITY;
  u32 i = kNoIdx;
  inline void take(float cv, u32 ci) {
#if ORACLE_MUTANT == 6
    if (cv <= v) {  // planted: last equal candidate wins
#else
    if (cv < v || (cv == v && ci < i)) {
#endif
      v = cv;
      i = ci;
    }
  }
};

struct Levels {
  u32 num_levels;
  u64 size[kMaxLevels];
  u64 off[kMaxLevels];
};

// src/xxx.h:225-243
u64 plan(u64 n, u64 t, u32 R_log, Levels* L) {
  const u64 R = 1ull << R_log;
  u64 size = n;
  L->num_levels = 1;
  L->off[0] = 0xFFFFFFFFull;  // ~uint32_t(0), src/xxx.h:229
  L->size[0] = size;
  u64 offset = 0;
  while (true) {
    if (L->num_levels >= kMaxLevels || size <= t) break;
    size = (size + R - 1) >> R_log;
#if ORACLE_MUTANT != 3  // planted: no 32-element alignment
    offset = (offset + kAlignment - 1) & ~(kAlignment - 1);
#endif
    L->off[L->num_levels] = offset;
    L->size[L->num_levels] = size;
    L->num_levels++;
    offset += size;
  }
  return offset;
}

struct View {
  const float* A;
  const float* lv;
  const u32* li;
  const Levels* L;
  inline float val(u32 level, u64 pos) const { return level == 0 ? A[pos] : lv[L->off[level] + pos]; }
  inline u32 idx(u32 level, u64 pos) const {
    return level == 0 ? static_cast<u32>(pos) : li[L->off[level] + pos];
  }
};

// Visit every level-k position in [lo, hi) as a candidate.
template <typename F>
inline void visit(u32 level, u64 lo, u64 hi, F&& f) {
  for (u64 p = lo; p < hi; ++p) f(level, p);
}

// Traversal of algs 16 and 24 (src/xxx.h:158-187 and 493-521): r is exclusive; while the
// remaining range is longer than 2R, take the partial blocks [l, ceil_R(l)) and
// [floor_R(r), r) at this level and move up; at the end scan [l, r) at the current level.
template <typename F>
void traverse_cl(u32 R_log, const Levels& L, u32 l, u32 r_incl, F&& f) {
  const u64 R = 1ull << R_log;
  u64 cl = l, cr = static_cast<u64>(r_incl) + 1;
  u32 level = 0;
  for (; level + 1 < L.num_levels; ++level) {
    if (cr - cl <= 2 * R) break;
    const u64 next_left = (cl + R - 1) & ~(R - 1);
    const u64 prev_right = cr & ~(R - 1);
    // xxx_static_scan_min scans the aligned chunk and masks to [i_left, e_right)
    visit(level, cl, next_left, f);
#if ORACLE_MUTANT != 5  // planted: right partial block dropped
    visit(level, prev_right, cr, f);
#endif
    cl = next_left >> R_log;
    cr = prev_right >> R_log;
  }
  visit(level, cl, cr, f);  // xxx_scan_min, src/xxx.h:185-186
}

// Traversal of alg 20 (src/xxx.h:899-927). R = 4 * 2^c. Break when r - l <= R or when the
// aligned bounds cross. vec_scan_min_left covers [l, align_down(l) + R) — the whole chunk when
// l is aligned — and vec_scan_min_right covers [floor_R(r), r) (src/xxx.h:794-856).
template <typename F>
void traverse_vl(u32 R_log, const Levels& L, u32 l, u32 r_incl, F&& f) {
  const u64 R = 1ull << R_log;
  u64 cl = l, cr = static_cast<u64>(r_incl) + 1;
  u32 level = 0;
  for (; level + 1 < L.num_levels; ++level) {
    if (cr - cl <= R) break;
    const u64 next_left = (cl + R - 1) & ~(R - 1);
    const u64 prev_right = cr & ~(R - 1);
    if (prev_right < next_left) break;
    const u64 start = cl & ~(R - 1);
    visit(level, cl, start + R, f);
    visit(level, prev_right, cr, f);
    cl = next_left >> R_log;
    cr = prev_right >> R_log;
  }
  visit(level, cl, cr, f);  // vec_scan_min, src/xxx.h:858-880
}

template <typename F>
void traverse(int kind, u32 R_log, const Levels& L, u32 l, u32 r, F&& f) {
  if (kind == 20) {
    traverse_vl(R_log, L, l, r, f);
  } else {
    traverse_cl(R_log, L, l, r, f);
  }
}

bool levels_from_args(u32 num_levels, const u64* sizes, const u64* offsets, Levels* L) {
  if (num_levels < 1 || num_levels > kMaxLevels) return false;
  L->num_levels = num_levels;
  for (u32 k = 0; k < num_levels; ++k) {
    L->size[k] = sizes[k];
    L->off[k] = offsets[k];
  }
  return true;
}

}  // namespace

extern "C" {

// Threads an OpenMP parallel region gets here (the CPU baseline reports it as its core count).
int oracle_num_threads() { return omp_get_max_threads(); }

// alg 0: value of the minimum (src/cpu_methods.h:18-39); OpenMP parallel for over queries.
void oracle_cpu_base_value(const float* A, u64 n, const u32* lr, u64 q, float* out, int nthreads) {
  (void)n;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < static_cast<int64_t>(q); ++i) {
    out[i] = cpu_min(A, lr[2 * i], lr[2 * i + 1]);
  }
}

// alg 100: leftmost index of the minimum (src/cpu_methods.h:55-77).
void oracle_cpu_base_index(const float* A, u64 n, const u32* lr, u64 q, u32* out, int nthreads) {
  (void)n;
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < static_cast<int64_t>(q); ++i) {
    out[i] = cpu_min_idx(A, lr[2 * i], lr[2 * i + 1]);
  }
}

// Level metadata; returns the level-buffer length in floats (src/xxx.h:225-243).
u64 oracle_plan(u64 n, u64 t, u32 R_log, u32* num_levels, u64* sizes, u64* offsets) {
  Levels L;
  const u64 floats = plan(n, t, R_log, &L);
  *num_levels = L.num_levels;
  for (u32 k = 0; k < L.num_levels; ++k) {
    sizes[k] = L.size[k];
    offsets[k] = L.off[k];
  }
  return floats;
}

// Level contents: level k+1 entry j = min of level k over [j*R, min(j*R+R, size_k))
// (src/xxx.h:105-118), with the leftmost level-0 index of that minimum.
int oracle_build(const float* A, u32 R_log, u32 num_levels, const u64* sizes, const u64* offsets,
                 float* lvl_val, u32* lvl_idx) {
  Levels L;
  if (!levels_from_args(num_levels, sizes, offsets, &L)) return -1;
  const u64 R = 1ull << R_log;
  for (u32 k = 0; k + 1 < L.num_levels; ++k) {
    const u64 src_size = L.size[k];
    const u64 dst_size = L.size[k + 1];
#pragma omp parallel for schedule(static)
    for (int64_t j = 0; j < static_cast<int64_t>(dst_size); ++j) {
      Best b;
      const u64 lo = static_cast<u64>(j) * R;
#if ORACLE_MUTANT == 4
      const u64 hi = lo + R - 1 < src_size ? lo + R - 1 : src_size;  // planted: short block
#else
      const u64 hi = lo + R < src_size ? lo + R : src_size;
#endif
      for (u64 p = lo; p < hi; ++p) {
        const float v = k == 0 ? A[p] : lvl_val[L.off[k] + p];
        const u32 ix = k == 0 ? static_cast<u32>(p) : lvl_idx[L.off[k] + p];
        b.take(v, ix);
      }
      lvl_val[L.off[k + 1] + j] = b.v;
      lvl_idx[L.off[k + 1] + j] = b.i;
    }
  }
  return 0;
}

// Hierarchical query on the CPU: kind 16/24 (CL traversal) or 20 (VL traversal).
// out_val[i] = A[leftmost argmin] (bit pattern), out_idx[i] = leftmost argmin; either may be NULL.
int oracle_hier_query(int kind, u32 R_log, const float* A, u32 num_levels, const u64* sizes,
                      const u64* offsets, const float* lvl_val, const u32* lvl_idx, const u32* lr,
                      u64 q, float* out_val, u32* out_idx, int nthreads) {
  Levels L;
  if (!levels_from_args(num_levels, sizes, offsets, &L)) return -1;
  const View V{A, lvl_val, lvl_idx, &L};
  if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 64)
  for (int64_t i = 0; i < static_cast<int64_t>(q); ++i) {
    Best b;
    traverse(kind, R_log, L, lr[2 * i], lr[2 * i + 1],
             [&](u32 level, u64 p) { b.take(V.val(level, p), V.idx(level, p)); });
    if (out_idx) out_idx[i] = b.i;
    if (out_val) out_val[i] = A[b.i];
  }
  return 0;
}

// Coverage of one query's traversal: cover[x] += 1 for every level-0 element x that a
// visited piece stands for (an entry at level k stands for R^k elements, clipped to n).
// The union must be exactly [l, r]; for kinds 16/24 every element is covered once.
int oracle_cover(int kind, u32 R_log, u32 num_levels, const u64* sizes, const u64* offsets, u32 l,
                 u32 r, u32* cover) {
  Levels L;
  if (!levels_from_args(num_levels, sizes, offsets, &L)) return -1;
  const u64 n = L.size[0];
  traverse(kind, R_log, L, l, r, [&](u32 level, u64 p) {
    const u64 span = 1ull << (R_log * level);
    const u64 lo = p * span;
    u64 hi = lo + span;
    if (hi > n) hi = n;
    for (u64 x = lo; x < hi; ++x) cover[x] += 1;
  });
  return 0;
}

}  // extern "C"
