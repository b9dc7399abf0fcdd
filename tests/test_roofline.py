"""The byte models behind the bench line's roofline figures (rmq_amd.roofline), on the CPU.

level0_sectors: the 64-byte sectors of level 0 the reference traversal reads (src/xxx.h:163-176,
185-186), against a brute-force set of touched sectors. build_bytes: level 0 read once plus
every upper-level entry and offset entry written once, against the library's own level plan
(rmq_plan, the metadata recurrence of src/xxx.h:225-243)."""

import numpy as np

import rmq_amd
from rmq_amd import roofline


def _brute_sectors(l, r):
    e = r + 1
    R0 = roofline.R0
    if e - l <= 2 * R0:
        idx = range(l, e)
    else:
        nl = -(-l // R0) * R0
        pr = (e // R0) * R0
        idx = list(range(l, nl)) + list(range(pr, e))
    return len({i // 16 for i in idx})


def test_level0_sectors_brute_force():
    rng = np.random.default_rng(3)
    n = 5000
    l = rng.integers(0, n, 3000)
    r = np.minimum(n - 1, l + np.exp2(rng.uniform(0, 11, 3000)).astype(np.int64))
    lr = np.stack([l, r], axis=1).astype(np.uint32)
    got = roofline.level0_sectors(lr)
    want = np.array([_brute_sectors(int(a), int(b)) for a, b in lr])
    assert np.array_equal(got, want)


def _rel(span_log):
    return 1 if span_log <= 8 else (2 if span_log <= 16 else 4)


def test_build_bytes_matches_level_plan():
    rng = np.random.default_rng(4)
    cases = [(1 << 26, 8192, 20, dict(cg_amount_log=0)), (1 << 26, 1 << 15, 16, {}),
             (1000, 64, 16, {}), (50, 64, 16, {}), (12345678, 64, 20, dict(cg_amount_log=1))]
    for _ in range(20):
        cases.append((int(rng.integers(1, 1 << 24)), 64, 16, {}))
    ns, ts, rls, want = [], [], [], 0
    for n, t, alg, opts in cases:
        p = rmq_amd.plan(n, t, alg, **opts)
        rl = int(p.R).bit_length() - 1
        one = 4 * n + sum(p.sizes[k] * (4 + _rel(rl * k)) for k in range(1, p.num_levels))
        assert roofline.build_bytes(n, t, rl) == one, (n, t, alg, opts)
        ns.append(n)
        ts.append(t)
        rls.append(rl)
        want += one
    assert roofline.build_bytes(ns, ts, rls) == want
