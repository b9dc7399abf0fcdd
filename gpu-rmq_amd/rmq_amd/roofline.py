"""Algorithmic bytes per query and the HBM roofline of the query kernel.

The figure is fixed by the reference traversal at R0 = 32 (algs 16/24 at both the default
g*c and the paper g*c), independent of the R this library runs with:
  e = r + 1; sec(a, b) = floor(b/16) - floor(a/16) + 1 if b >= a else 0   (64-byte sectors)
  if e - l <= 2*R0: S0 = sec(l, r)
  else:             S0 = sec(l, ceil32(l) - 1) + sec(floor32(e), e - 1)
  B_q = 8 (the {l, r} pair) + 4 * k_out (value or index: 1, both: 2) + 64 * S0
These are the level-0 pieces the reference reads (src/xxx.h:163-176, 185-186). Upper levels
are <= n/32 floats and are left out: they stay in the 256 MB Infinity Cache / L2.
"""

from __future__ import annotations

import numpy as np

R0 = 32
HBM_PEAK_GBPS = 8000.0  # MI355X HBM3E

# Instruction issue of MI355X: 256 CUs at 2.4 GHz. A CU has four SIMDs of 16 lanes; a wave64
# VALU instruction holds its SIMD for 4 cycles and the sequencer issues at most one VALU and one
# SALU instruction per CU per cycle, so both peaks are one wave-instruction per CU per cycle.
# The counters SQ_INSTS_VALU / SQ_INSTS_SALU count wave-instructions (one per wavefront that
# issues it, whatever its active lanes), so per-dispatch counts divided by the kernel time are
# directly comparable with these peaks.
CUS = 256
CLOCK_GHZ = 2.4
VALU_PEAK_GINST = CUS * CLOCK_GHZ  # 614.4 G wave-instructions/s
SALU_PEAK_GINST = CUS * CLOCK_GHZ
# What gpu-rmq_amd/issue_probe measured on an MI355X (G wave-instructions/s, 8 waves per SIMD,
# independent chains): 32-bit integer adds issue faster than the one-per-CU-cycle model, float
# min and selects at it. bench.py measures them live; these are for the profile summaries.
VALU_PEAK_FASTEST = 983.1  # v_add_u32, 1.60 per CU per cycle
VALU_PEAK_FP32 = 599.6     # v_min_f32 (v_cndmask_b32: 596.3), 0.98 per CU per cycle
SALU_PEAK_MEASURED = 591.9  # s_mov_b32


def _sec(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    s = (b // 16) - (a // 16) + 1
    return np.where(b >= a, s, 0)


def level0_sectors(lr: np.ndarray) -> np.ndarray:
    l = lr[:, 0].astype(np.int64)
    r = lr[:, 1].astype(np.int64)
    e = r + 1
    short = (e - l) <= 2 * R0
    nl = ((l + R0 - 1) // R0) * R0
    pr = (e // R0) * R0
    long_s = _sec(l, nl - 1) + _sec(pr, e - 1)
    return np.where(short, _sec(l, r), long_s)


def bytes_per_query(lr: np.ndarray, k_out: int) -> np.ndarray:
    return 8 + 4 * k_out + 64 * level0_sectors(lr)


def total_bytes(lr: np.ndarray, k_out: int, chunk: int = 1 << 22) -> int:
    tot = 0
    for s in range(0, lr.shape[0], chunk):
        tot += int(bytes_per_query(lr[s:s + chunk], k_out).sum())
    return tot


def _rel_bytes(span_log: np.ndarray) -> np.ndarray:
    """Bytes of one offset entry whose block spans 2^span_log level-0 elements."""
    return np.where(span_log <= 8, 1, np.where(span_log <= 16, 2, 4))


def build_bytes(n, t, r_log) -> int:
    """Algorithmic HBM bytes of building hierarchies over arrays of n elements (scalars or
    arrays, one per hierarchy) with scan threshold t and R = 2^r_log: level 0 read once (4 B per
    element), every upper-level entry written once (4 B value + its offset entry). Levels follow
    the metadata recurrence (src/xxx.h:225-243): while a level holds more than t entries and
    fewer than 20 levels exist, the next holds ceil(size / R)."""
    n = np.atleast_1d(np.asarray(n, dtype=np.int64))
    t = np.broadcast_to(np.asarray(t, dtype=np.int64), n.shape)
    rl = np.broadcast_to(np.asarray(r_log, dtype=np.int64), n.shape)
    total = 4 * int(n.sum())
    size = n.copy()
    for k in range(1, 20):
        live = size > t
        if not live.any():
            break
        size = np.where(live, (size + (1 << rl) - 1) >> rl, size)
        total += int(np.where(live, size * (4 + _rel_bytes(rl * k)), 0).sum())
    return total
