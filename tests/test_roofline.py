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


def test_bench_bound_skips_passed_measured_ceilings(monkeypatch):
    """bench.roofline_block: the bound is the resource nearest its peak, except a measured HBM
    ceiling that the kernel passes (misses the Infinity Cache serves count in FETCH_SIZE but not
    in the probes over 4 GiB buffers): that one is listed and skipped. Figures of the n = 2^30
    small-lengths kernel: 22.3 GB per launch of 3.06 ms."""
    import bench

    counters = {"bytes_per_launch": 22_295_000_000, "fetch_bytes": 22_027_000_000,
                "write_bytes": 268_000_000, "valu_insts": 545_300_000, "salu_insts": 253_500_000}
    monkeypatch.setattr(bench, "measured_counters", lambda key: counters)
    ceilings = {"read_gbps": 5972.1, "read_nt_gbps": 6456.0, "copy_gbps": 5176.4,
                "read4w1_gbps": 5255.7, "random128_gsegments_per_s": 48.12}
    peaks = {"valu_peak_ginst_per_s": 1005.6, "salu_peak_ginst_per_s": 592.2}
    out = bench.roofline_block("k", 196.0 * 2**26, 2**26, 3.06, peaks, ceilings)
    assert out["passed_measured_ceilings"] == ["hbm_achievable", "hbm_random_lines"]
    assert out["bound"] == "hbm" and out["peak"] == roofline.HBM_PEAK_GBPS
    assert abs(out["frac"] - 22_295_000_000 / 3.06e-3 / 1e9 / 8000.0) < 1e-4
    assert out["resources"]["hbm_random_lines"]["frac"] > 1
    # a kernel below every ceiling: the measured stream ceiling binds (0.62 > 0.50 of 8 TB/s)
    counters.update(bytes_per_launch=1_033_000_000, fetch_bytes=715_000_000, valu_insts=103_200_000,
                    salu_insts=58_800_000)
    out = bench.roofline_block("k", 196.0 * 2**26, 2**26, 0.2595, peaks, ceilings)
    assert "passed_measured_ceilings" not in out
    assert out["bound"] == "hbm_achievable" and out["frac"] <= 1
    # nothing measured for the key: unmeasured, no fraction
    monkeypatch.setattr(bench, "measured_counters", lambda key: None)
    out = bench.roofline_block("k", 196.0 * 2**26, 2**26, 1.0, peaks, ceilings)
    assert out["bound"] == "unmeasured" and out["frac"] is None
