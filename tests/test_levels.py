"""Level contents read back from the device (rmq_read_levels, rmq_batch_read_levels) against the
oracle's level build (oracle_build: xxx_build_step, src/xxx.h:91-120, plus the leftmost level-0
index of every block minimum), bit-exactly: every entry's value bit pattern and index, over the
whole level buffer in the reference's layout (src/xxx.h:225-249), alignment gaps included.

The query tests see the levels only through answers; these pin the build itself, for every
reduction factor the plain build offers, ragged sizes, ties, signed zeros, +inf and NaN runs, the
headline shape (n = 2^26, R = 4, t = 2^13) and the reference default (alg 16, t = 2^15), and for
the batched build (one launch per level over many arrays of mixed reduction factors)."""

import sys

import numpy as np
import pytest

import oracle
import rmq_amd

pytestmark = pytest.mark.gpu

NO_IDX = 0xFFFFFFFF

# (alg, build options, R_log, t)
PLAIN = [
    (16, {}, 5, 64),
    (16, {}, 5, 1 << 15),
    (24, {}, 5, 64),
    (20, {}, 4, 16),
    (20, dict(cg_amount_log=0), 2, 4),
    (20, dict(cg_amount_log=1), 3, 8),
    (20, dict(cg_amount_log=4), 6, 64),
    (16, dict(cg_size_log=3, cg_amount_log=0), 3, 16),
    (16, dict(cg_size_log=3, cg_amount_log=3), 6, 128),
    (24, dict(cg_size_log=4, cg_amount_log=3, group_lanes=64), 7, 256),
]

SIZES = [1, 5, 33, 1000, 4097, (1 << 16) + 3, (1 << 20) + 17]


def _say(msg):
    sys.stderr.write(f"[gpu-case] {msg}\n")
    sys.stderr.flush()


def _array(kind, n, rng):
    if kind == "uniform":
        return rng.random(n, dtype=np.float32)
    if kind == "ties":
        return rng.integers(0, 8, n).astype(np.float32)
    if kind == "zeros":
        A = np.where(rng.random(n) < 0.5, np.float32(0.0), np.float32(-0.0)).astype(np.float32)
        A[rng.random(n) < 0.05] = 1.0
        return A
    if kind == "inf":
        A = np.full(n, np.inf, dtype=np.float32)
        A[rng.random(n) < 0.01] = 2.0
        return A
    if kind == "desc":
        return np.arange(n, 0, -1, dtype=np.float32)
    if kind == "nan":
        A = rng.random(n, dtype=np.float32)
        A[rng.random(n) < 0.3] = np.nan
        A[: min(n, 300)] = np.nan  # whole blocks of NaN
        return A
    raise ValueError(kind)


def _compare(got_v, got_i, ref: "oracle.CpuHierarchy", what):
    gv = got_v.cpu().numpy().view(np.uint32)
    gi = got_i.cpu().numpy().view(np.uint32)
    rv = ref.lv[: ref.floats].view(np.uint32)
    ri = ref.li[: ref.floats]
    assert gv.shape == rv.shape, what
    assert np.array_equal(gv, rv), (what, int(np.flatnonzero(gv != rv)[0]))
    known = ri != NO_IDX
    assert np.array_equal(gi[known], ri[known]), (what, int(np.flatnonzero((gi != ri) & known)[0]))
    # a block of nothing but NaN: the oracle keeps no index; the device keeps the block's first
    # level-0 position with value +inf, which no query can pick over a real element
    assert (ref.lv[: ref.floats][~known] == np.inf).all(), what


@pytest.mark.parametrize("alg,opts,R_log,t", PLAIN)
@pytest.mark.parametrize("kind", ["uniform", "ties", "zeros", "inf", "desc", "nan"])
def test_plain_levels_match_oracle(device, alg, opts, R_log, t, kind):
    import torch

    rng = np.random.default_rng(R_log * 100 + len(kind))
    for n in SIZES:
        A = _array(kind, n, rng)
        _say(f"levels alg={alg} {opts} t={t} n={n} {kind}")
        dA = torch.from_numpy(A).to(device)
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            md = h.metadata()
            assert md.R == 1 << R_log
            ref = oracle.CpuHierarchy(A, t, R_log)
            assert md.num_levels == ref.num_levels and md.buffer_floats == ref.floats
            v, i = h.levels()
            torch.cuda.synchronize()
        _compare(v, i, ref, (alg, opts, t, n, kind))


@pytest.mark.parametrize("alg,opts,R_log,t", [(20, dict(cg_amount_log=0), 2, 8192), (16, {}, 5, 1 << 15)])
def test_full_size_levels(device, alg, opts, R_log, t):
    """n = 2^26 (the headline array), from the reference generator's seed."""
    import torch

    n = 1 << 26
    A = rmq_amd.gen_array(n, 27722)
    _say(f"levels full size alg={alg} {opts} t={t}")
    dA = torch.from_numpy(A).to(device)
    with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
        v, i = h.levels()
        torch.cuda.synchronize()
    _compare(v, i, oracle.CpuHierarchy(A, t, R_log), (alg, opts, t))


def test_levels_of_one_level_hierarchy_and_bad_arguments(device):
    import ctypes

    import torch

    dA = torch.rand(50, device=device)
    with rmq_amd.Hierarchy(dA, 64) as h:  # n <= t: no level buffer
        v, i = h.levels()
        assert v.numel() == 0 and i.numel() == 0
        L = rmq_amd.lib()
        assert L.rmq_read_levels(h._h, None, None, None) == rmq_amd.RMQ_EINVAL
    assert rmq_amd.lib().rmq_read_levels(None, ctypes.c_void_p(16), None, None) == rmq_amd.RMQ_EINVAL


RLOG_OF = {"lane": 5, "lane8": 3, "lane4": 2, "group4": 5, "group8": 5, "group16": 5,
           "group32": 5, "wave": 5}


def test_batched_levels_match_oracle(device):
    """One batch holding every variant (R = 4, 8 and 32 builds in the same launches), sizes
    log-U[2^4, 2^19], thresholds from each variant's minimum up to 2^12, ties planted."""
    import torch

    rng = np.random.default_rng(99)
    names = list(RLOG_OF)
    count = 120
    sizes = np.exp2(rng.uniform(4, 19, count)).astype(np.int64)
    offsets = np.zeros(count, dtype=np.int64)
    offsets[1:] = np.cumsum((sizes[:-1] + 63) // 64 * 64)
    var = [names[a % len(names)] for a in range(count)]
    ts = []
    for a in range(count):
        lo = 1 << RLOG_OF[var[a]]
        lo = lo if var[a] in ("lane4", "lane8") else 64
        ts.append(int(rng.choice([lo, 2 * lo, 256, 4096])))
    host = rng.random(int(offsets[-1] + sizes[-1] + 64), dtype=np.float32)
    for a in range(0, count, 5):
        o, n = int(offsets[a]), int(sizes[a])
        host[o:o + n] = np.floor(host[o:o + n] * 4)  # ties
    dA = torch.from_numpy(host).to(device)
    _say(f"batched levels: {count} arrays")
    with rmq_amd.Batch(dA, offsets, sizes, ts, var) as b:
        for a in range(count):
            o, n = int(offsets[a]), int(sizes[a])
            ref = oracle.CpuHierarchy(host[o:o + n], ts[a], RLOG_OF[var[a]])
            v, i = b.levels(a, ref.floats)
            torch.cuda.synchronize()
            _compare(v, i, ref, (a, var[a], ts[a], n))
