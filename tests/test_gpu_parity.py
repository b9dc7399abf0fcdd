"""GPU parity: librmq.so's HIP kernels against the CPU oracle, bit-exact (value bit pattern and
leftmost index), through the C ABI. Inputs come from the reference generator restatement.

Sizes: every test compares the full query set against the oracle where the oracle finishes in
seconds (CPU BASE for short ranges, the oracle's CPU hierarchy — itself pinned equal to CPU BASE
in test_oracle.py — for long ones) plus a CPU BASE sample; the full-size runs add
size-independent properties (index inside [l, r], value == A[index], all algorithms agree).
"""

import sys

import numpy as np
import pytest

import oracle
import rmq_amd

pytestmark = pytest.mark.gpu

SEED = 27722  # SEEDS[0], scripts/perf-benchmark-exp.sh:20

# (alg, extra build options): reference defaults and paper settings plus execution variants
VARIANTS = [
    (16, {}),
    (24, {}),
    (20, {}),
    (16, dict(cg_size_log=4, cg_amount_log=1)),             # paper CL: g=16, c=2 -> R=32
    (20, dict(cg_amount_log=1)),                             # paper VL: R=8
    (16, dict(group_lanes=4)),
    (16, dict(group_lanes=16)),
    (24, dict(group_lanes=16)),
    (16, dict(group_lanes=64)),                              # a wave per query
    (16, dict(group_lanes=1)),                               # a lane per query, R=32
    (24, dict(group_lanes=1, cg_size_log=3, cg_amount_log=3)),  # a lane per query, R=64
    (24, dict(group_lanes=32)),
    (24, dict(group_lanes=64, cg_size_log=4, cg_amount_log=3)),  # a wave per query, R=128
    (16, dict(cg_size_log=3, cg_amount_log=0)),              # R=8
    (16, dict(cg_size_log=3, cg_amount_log=3)),              # R=64
    (24, dict(cg_size_log=4, cg_amount_log=3)),              # R=128
    (20, dict(cg_amount_log=0)),                             # R=4
    (20, dict(cg_amount_log=4)),                             # R=64
    (16, dict(full=True)),                                 # level loop instead of up-front loads
    (24, dict(full=True, group_lanes=4)),
    # descents packed onto full wavefronts (lane solver), and ignored by the other solvers
    (20, dict(cg_amount_log=0, pack=True)),
    (20, dict(cg_amount_log=1, pack=True)),
    (20, dict(cg_amount_log=4, pack=True)),
    (16, dict(pack=True)),
    (24, dict(group_lanes=1, cg_size_log=3, cg_amount_log=3, pack=True)),
    (16, dict(group_lanes=4, pack=True)),
    # two queries per lane (their top-level table reads in flight together)
    (20, dict(cg_amount_log=0, paired=True)),
    (20, dict(cg_amount_log=1, paired=True)),
    (16, dict(paired=True)),
    (20, dict(cg_amount_log=0, paired=True, pack=True)),   # two per lane, descents packed
    (16, dict(paired=True, pack=True)),
    # block prefix / suffix minima, read by the lane and group solvers where R >= 16 (built by
    # default at R >= 32, so the defaults above read them)
    (16, dict(group_lanes=8, scans=False)),
    (24, dict(group_lanes=4, scans=False)),
    (16, dict(group_lanes=64, scans=False)),
    (24, dict(group_lanes=4, cg_size_log=2, cg_amount_log=2, scans=True)),  # R = 16
    (16, dict(group_lanes=16, cg_size_log=4, cg_amount_log=2, scans=True)),  # R = 64
    (20, dict(cg_amount_log=2, scans=True)),
    (20, dict(cg_amount_log=2, scans=True, pack=True)),
    (20, dict(cg_amount_log=2, scans=True, paired=True)),
    (16, dict(scans=False)),
    (16, dict(paired=True, scans=False)),
    (20, dict(cg_amount_log=3, scans=False, pack=True)),
    (24, dict(group_lanes=1, cg_size_log=3, cg_amount_log=3, scans=False)),
    (16, dict(group_lanes=4, scans=True)),
]


def _say(msg):
    sys.stderr.write(f"[gpu-case] {msg}\n")
    sys.stderr.flush()


def _R(alg, opts):
    opts = {k: v for k, v in opts.items() if k in ("cg_size_log", "cg_amount_log")}
    g = opts.get("cg_size_log", 3)
    c = opts.get("cg_amount_log", 2)
    return 1 << (c + 2) if alg == 20 else 1 << (g + c)


def _min_t(alg, opts):
    return _R(alg, opts) * (1 if alg == 20 else 2)


def run_gpu(A, lr, alg, t, want_val=True, want_idx=True, **opts):
    import torch

    dA = torch.from_numpy(np.ascontiguousarray(A)).cuda()
    dQ = torch.from_numpy(np.ascontiguousarray(lr).view(np.int32)).cuda()
    with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
        v, i = h.query(dQ, want_val=want_val, want_idx=want_idx)
        torch.cuda.synchronize()
    v = v.cpu().numpy() if v is not None else None
    i = i.cpu().numpy().view(np.uint32) if i is not None else None
    return v, i


def assert_exact(A, lr, v, i, ref_i, what):
    if i is not None:
        bad = np.nonzero(i != ref_i)[0]
        assert bad.size == 0, f"{what}: {bad.size} index mismatches, first q={bad[:5]} " \
                              f"lr={lr[bad[:3]].tolist()} got={i[bad[:3]]} want={ref_i[bad[:3]]}"
    if v is not None:
        want = A[ref_i].view(np.uint32)
        bad = np.nonzero(v.view(np.uint32) != want)[0]
        assert bad.size == 0, f"{what}: {bad.size} value mismatches, first q={bad[:5]}"


def edge_queries(n, R):
    pairs = {(0, n - 1), (0, 0), (n - 1, n - 1)}
    for L in {1, 2, 3, 4, 5, R - 1, R, R + 1, 2 * R - 1, 2 * R, 2 * R + 1, 2 * R + 2, 3 * R,
              4 * R + 3, R * R, R * R + 1, 2 * R * R + 5}:
        if 1 <= L <= n:
            for l in {0, 1, 3, R - 1, R, R + 1, n - L, max(0, n - L - 1), (n - L) // 2}:
                if 0 <= l <= n - L:
                    pairs.add((l, l + L - 1))
    return np.array(sorted(pairs), dtype=np.uint32)


@pytest.mark.parametrize("alg,opts", VARIANTS, ids=lambda x: str(x))
def test_small_edge_cases(device, alg, opts):
    rng = np.random.default_rng(1)
    R = _R(alg, opts)
    tmin = _min_t(alg, opts)
    for n in (1, 2, 3, 5, 33, 1001, 4099, 20000, 65536 + 7):
        for t in sorted({tmin, 4 * tmin, 1 << 15}):
            for kind in ("uniform", "ties", "zeros"):
                if kind == "uniform":
                    A = rng.random(n, dtype=np.float32)
                elif kind == "ties":
                    A = rng.integers(0, 3, size=n).astype(np.float32)
                else:
                    A = rng.integers(0, 2, size=n).astype(np.float32)
                    A[rng.random(n) < 0.5] *= np.float32(-1.0)  # mixes +0 and -0
                rnd = np.sort(rng.integers(0, n, size=(300, 2)), axis=1).astype(np.uint32)
                lr = np.concatenate([edge_queries(n, R), rnd])
                ref_i = oracle.cpu_base_index(A, lr)
                _say(f"small alg={alg} opts={opts} n={n} t={t} kind={kind} q={len(lr)}")
                v, i = run_gpu(A, lr, alg, t, True, True, **opts)
                assert_exact(A, lr, v, i, ref_i, f"alg={alg} n={n} t={t} {kind} both")
                v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
                assert_exact(A, lr, v, None, ref_i, f"alg={alg} n={n} t={t} {kind} value")
                _, i = run_gpu(A, lr, alg, t, False, True, **opts)
                assert_exact(A, lr, None, i, ref_i, f"alg={alg} n={n} t={t} {kind} index")


@pytest.mark.parametrize("t", [64, 128, 4096, 1 << 17, 1 << 23])
def test_every_depth(device, t):
    """n = 2^22 + 3 with t from 64 (5 levels) to 2^23 (1 level): every compiled depth of the
    up-front form and the rolled loop, value and index, mixed range lengths."""
    n = (1 << 22) + 3
    A = rmq_amd.gen_array(n, 77)
    rep = A[::1000].copy()
    A[::997][: rep.size] = rep  # plant repeated values
    lr = rmq_amd.gen_queries(1 << 16, -6, n, 77)
    ref_v, ref_i = oracle.CpuHierarchy(A, 64, 5, 16).query(lr)
    for alg, opts in ((16, {}), (24, {}), (16, dict(full=True)), (16, dict(group_lanes=4)),
                      (24, dict(group_lanes=16)), (16, dict(group_lanes=64)), (16, dict(group_lanes=1)),
                      (20, {})):
        _say(f"depth t={t} alg={alg} opts={opts}")
        v, i = run_gpu(A, lr, alg, t, True, True, **opts)
        assert_exact(A, lr, v, i, ref_i, f"depth t={t} alg={alg} {opts}")
        v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
        assert_exact(A, lr, v, None, ref_i, f"depth t={t} alg={alg} {opts} value")


@pytest.mark.parametrize("alg,opts", [(16, {}), (16, dict(group_lanes=1)), (24, dict(group_lanes=4)),
                                      (16, dict(group_lanes=64)), (20, dict(cg_amount_log=0)),
                                      (20, dict(cg_amount_log=1))], ids=str)
def test_top_table(device, alg, opts):
    """Large t leaves long final ranges on the top level, which the sparse table over it answers:
    the reference default t = 2^15 with uniform lengths, a one-level hierarchy (t >= n, the table
    over the array itself) and a small top level (about 100 entries). Values are
    quantised (many ties) with both zero signs planted; answers are the same with and without
    the table and equal to the oracle's."""
    rng = np.random.default_rng(alg * 7 + len(opts))
    R = _R(alg, opts)
    for n, t in ((1 << 20, 1 << 15), ((1 << 18) + 3, 1 << 19), (100 * R * R + 17, 128)):
        A = np.floor(rmq_amd.gen_array(n, 11) * 512.0).astype(np.float32)
        A[rng.integers(0, n, size=64)] = np.float32(-0.0)  # next to the +0 of A < 1/512
        lr = np.concatenate([rmq_amd.gen_queries(1 << 14, -1, n, 5), rmq_amd.gen_queries(1 << 12, -6, n, 6),
                             edge_queries(n, R)])
        ref_v, ref_i = oracle.CpuHierarchy(A, 64, 5, 16).query(lr)
        levels = rmq_amd.plan(n, t, alg, **{k: v for k, v in opts.items() if k != "group_lanes"})
        _say(f"top table alg={alg} {opts} n={n} t={t} top={levels.sizes[-1]}")
        for table in (True, False):
            v, i = run_gpu(A, lr, alg, t, True, True, top_table=table, **opts)
            assert_exact(A, lr, v, i, ref_i, f"alg={alg} {opts} n={n} t={t} table={table}")
            v, _ = run_gpu(A, lr, alg, t, True, False, top_table=table, **opts)
            assert_exact(A, lr, v, None, ref_i, f"alg={alg} {opts} n={n} t={t} table={table} value")


def test_tie_vectors(device):
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tie_vectors.json")
    with open(path) as f:
        cases = json.load(f)
    for case in cases:
        A = np.array(case["array_bits"], dtype=np.uint32).view(np.float32)
        lr = np.array(case["lr"], dtype=np.uint32)
        for alg, opts in VARIANTS:
            t = _min_t(alg, opts)
            _say(f"tie_vectors {case['name']} alg={alg} opts={opts}")
            v, i = run_gpu(A, lr, alg, t, True, True, **opts)
            assert i.tolist() == case["index"], (case["name"], alg, opts)
            assert v.view(np.uint32).tolist() == case["value_bits"], (case["name"], alg, opts)
            v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
            assert v.view(np.uint32).tolist() == case["value_bits"], (case["name"], alg, opts)


@pytest.mark.parametrize("signs", ["pos", "neg", "both", "none"])
def test_zero_signs_value_mode(device, signs):
    """Value-only queries whose minimum is zero: the zero's sign comes from the array's recorded
    zero signs, and only an array holding both +0 and -0 needs index tracking. Covers every
    variant, hierarchies of one level (t >= n: no level-0 pass records the signs) and many."""
    rng = np.random.default_rng({"pos": 1, "neg": 2, "both": 3, "none": 4}[signs])
    n = 20000
    A = rng.random(n, dtype=np.float32) + np.float32(0.5)
    zpos = rng.choice(n, 40, replace=False)
    if signs in ("pos", "both"):
        A[zpos[:20]] = np.float32(0.0)
    if signs in ("neg", "both"):
        A[zpos[20:]] = np.float32(-0.0)
    lr = np.sort(rng.integers(0, n, size=(4000, 2)), axis=1).astype(np.uint32)
    ref_i = oracle.cpu_base_index(A, lr)
    for alg, opts in VARIANTS:
        for t in (_min_t(alg, opts), 1 << 15):
            _say(f"zero_signs {signs} alg={alg} opts={opts} t={t}")
            v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
            assert_exact(A, lr, v, None, ref_i, f"zero signs {signs} alg={alg} {opts} t={t}")


def test_gpu_base_matches_cpu_base(device):
    rng = np.random.default_rng(4)
    import torch

    for n in (1, 77, 5000):
        A = rng.integers(0, 5, size=n).astype(np.float32)
        lr = np.sort(rng.integers(0, n, size=(2000, 2)), axis=1).astype(np.uint32)
        _say(f"gpu_base n={n}")
        v, i = rmq_amd.gpu_base(torch.from_numpy(A).cuda(), torch.from_numpy(lr.view(np.int32)).cuda(),
                                True, True)
        torch.cuda.synchronize()
        ref_i = oracle.cpu_base_index(A, lr)
        assert np.array_equal(i.cpu().numpy().view(np.uint32), ref_i)
        assert np.array_equal(v.cpu().numpy().view(np.uint32), A[ref_i].view(np.uint32))


def test_config1_plain_value(device):
    """n = 2^20, q = 2^20, lr = -1 (uniform large), t = 2^15, value mode."""
    n = q = 1 << 20
    A = rmq_amd.gen_array(n, SEED)
    lr = rmq_amd.gen_queries(q, -1, n, SEED)
    ref_v, ref_i = oracle.CpuHierarchy(A, 1 << 15, 5, 16).query(lr)
    sample = rmq_amd.gen_sample_indices(q, 1 << 12, SEED)
    assert np.array_equal(oracle.cpu_base_index(A, lr[sample]), ref_i[sample])
    for alg, opts in VARIANTS[:5] + [(20, dict(cg_amount_log=1, pack=True))]:
        _say(f"config1 alg={alg} opts={opts}")
        v, _ = run_gpu(A, lr, alg, 1 << 15, True, False, **opts)
        assert_exact(A, lr, v, None, ref_i, f"config1 alg={alg} {opts}")


@pytest.mark.parametrize("t_log", [10, 15, 18])
def test_config2_index_sampled_full_n(device, t_log):
    """n = 2^26, lr = -2 (lognormal medium), index mode, t swept; 2^22 queries vs the oracle."""
    n = 1 << 26
    A = _array_2_26()
    lr = _queries_2_26()[: 1 << 22]
    ref_v, ref_i = oracle.CpuHierarchy(A, 1 << 15, 5, 16).query(lr)
    for alg, opts in ((16, {}), (24, {}), (20, {}), (20, dict(cg_amount_log=0, pack=True)),
                      (20, dict(cg_amount_log=2, scans=True)), (16, dict(paired=True)), (16, dict(scans=False)),
                      # the cooperative forms, with block scans at R = 32
                      (16, dict(group_lanes=4)), (16, dict(group_lanes=8)), (24, dict(group_lanes=4))):
        _say(f"config2 alg={alg} opts={opts} t=2^{t_log}")
        v, i = run_gpu(A, lr, alg, 1 << t_log, True, True, **opts)
        assert_exact(A, lr, v, i, ref_i, f"config2 alg={alg} t=2^{t_log}")
    sample = rmq_amd.gen_sample_indices(lr.shape[0], 1 << 13, SEED)
    assert np.array_equal(oracle.cpu_base_index(A, lr[sample]), ref_i[sample])
    assert n == A.size


def test_config2_full_q_properties(device):
    """Full n = q = 2^26: every algorithm gives the same answer; answers are inside [l, r]."""
    import torch

    A = _array_2_26()
    lr = _queries_2_26()
    dA = torch.from_numpy(A).cuda()
    dQ = torch.from_numpy(lr.view(np.int32)).cuda()
    outs = {}
    # the last: the timed headline configuration (R = 4, t = 2^13, packed descents)
    for key, alg, t, opts in ((16, 16, 1 << 15, {}), (24, 24, 1 << 15, {}), (20, 20, 1 << 15, {}),
                              ("20p", 20, 1 << 13, dict(cg_amount_log=0, pack=True))):
        _say(f"config2 full q alg={alg} t={t} opts={opts}")
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            v, i = h.query(dQ, True, True)
            torch.cuda.synchronize()
        outs[key] = (v.cpu().numpy(), i.cpu().numpy().view(np.uint32))
    v16, i16 = outs[16]
    for key in (24, 20, "20p"):
        assert np.array_equal(outs[key][1], i16), key
        assert np.array_equal(outs[key][0].view(np.uint32), v16.view(np.uint32)), key
    assert (i16 >= lr[:, 0]).all() and (i16 <= lr[:, 1]).all()
    assert np.array_equal(A[i16].view(np.uint32), v16.view(np.uint32))
    sample = rmq_amd.gen_sample_indices(lr.shape[0], 1 << 13, SEED + 1)
    assert np.array_equal(oracle.cpu_base_index(A, lr[sample]), i16[sample])


_CACHE = {}


def _array_2_26():
    if "A26" not in _CACHE:
        _CACHE["A26"] = rmq_amd.gen_array(1 << 26, SEED)
    return _CACHE["A26"]


def _queries_2_26():
    if "Q26" not in _CACHE:
        _CACHE["Q26"] = rmq_amd.gen_queries(1 << 26, -2, 1 << 26, SEED)
    return _CACHE["Q26"]


@pytest.mark.parametrize("lr_mode", [-3, -6])
def test_config3_large_array(device, lr_mode):
    """n = 2^30 (4 GiB), lr = -3 and -6, value + index; 2^20 queries vs the oracle hierarchy, in
    every form the bench times at this size among others."""
    import torch

    n = 1 << 30
    if "A30" not in _CACHE:
        _CACHE.clear()
        _CACHE["A30"] = rmq_amd.gen_array(n, SEED)
    A = _CACHE["A30"]
    lr = rmq_amd.gen_queries(1 << 20, lr_mode, n, SEED)
    ref_v, ref_i = oracle.CpuHierarchy(A, 64, 5, 16).query(lr)
    dA = torch.from_numpy(A).cuda()
    dQ = torch.from_numpy(lr.view(np.int32)).cuda()
    for alg, t, opts in ((16, 64, {}), (24, 1 << 15, {}), (20, 32, dict(cg_amount_log=1)),
                         (20, 1 << 13, dict(cg_amount_log=0, pack=True)),
                         (16, 1 << 10, dict(group_lanes=4)), (20, 64, dict(cg_amount_log=1)),
                         (20, 1 << 10, dict(cg_amount_log=1, paired=True)), (20, 64, dict(cg_amount_log=0))):
        _say(f"config3 lr={lr_mode} alg={alg} t={t} opts={opts}")
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            v, i = h.query(dQ, True, True)
            torch.cuda.synchronize()
        assert_exact(A, lr, v.cpu().numpy(), i.cpu().numpy().view(np.uint32), ref_i,
                     f"config3 lr={lr_mode} alg={alg}")
    del dA
    torch.cuda.empty_cache()


def test_array_at_int32_limit(device):
    """n = 2^31 - 1 (the reference's largest int32 n; 8 GiB), one piece: level-0 spans of upper
    entries reach 2^31, so the winner codes and span keys use every bit below the top one.
    Queries: mixed lengths and ranges ending at n - 1; short ones against GPU BASE, long ones
    against torch.argmin (first occurrence), every answer inside its range with value ==
    A[index]."""
    import torch

    _CACHE.clear()
    n = (1 << 31) - 1
    _say("int32 limit: fill")
    dA = torch.empty(n, dtype=torch.float32, device="cuda")
    rmq_amd.fill_uniform(dA, seed=17)  # 24-bit values: long ranges hold ties at their minimum
    lr = rmq_amd.gen_queries(1 << 16, -6, n, 17)
    edge = np.array([[0, n - 1], [n - 1, n - 1], [n - 5, n - 1], [n - 70000, n - 1], [1 << 30, n - 1],
                     [(1 << 30) - 3, (1 << 30) + 3], [n - (1 << 24), n - 2]], dtype=np.uint32)
    lr = np.concatenate([edge, lr])
    dQ = torch.from_numpy(lr.view(np.int32)).cuda()
    short = np.nonzero((lr[:, 1].astype(np.int64) - lr[:, 0]) < (1 << 20))[0]
    sq = torch.from_numpy(np.ascontiguousarray(lr[short]).view(np.int32)).cuda()
    _, base_i = rmq_amd.gpu_base(dA, sq, False, True)
    torch.cuda.synchronize()
    base_i = base_i.cpu().numpy().view(np.uint32)
    lo = torch.from_numpy(lr[:, 0].astype(np.int64)).cuda()
    hi = torch.from_numpy(lr[:, 1].astype(np.int64)).cuda()
    for alg, opts, t in ((20, dict(cg_amount_log=0), 8192), (16, {}, 1 << 15), (16, dict(group_lanes=4), 64),
                         (20, dict(cg_amount_log=1), 1 << 17), (20, dict(cg_amount_log=0, pack=True), 8192)):
        _say(f"int32 limit alg={alg} opts={opts} t={t}")
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            v, i = h.query(dQ, True, True)
            torch.cuda.synchronize()
        i64 = i.to(torch.int64) & 0xFFFFFFFF
        assert bool(((i64 >= lo) & (i64 <= hi)).all()), (alg, opts)
        assert torch.equal(v.view(torch.int32), dA[i64].view(torch.int32)), (alg, opts)
        ih = i.cpu().numpy().view(np.uint32)
        assert np.array_equal(ih[short], base_i), (alg, opts)
        for k in range(edge.shape[0]):
            l, r = int(lr[k, 0]), int(lr[k, 1])
            assert int(ih[k]) == l + int(torch.argmin(dA[l:r + 1])), (alg, opts, k, l, r)
    del dA
    torch.cuda.empty_cache()


def test_streams_and_reuse(device):
    """One resident hierarchy, queried concurrently on two streams and repeatedly."""
    import torch

    n = 1 << 22
    A = rmq_amd.gen_array(n, 5)
    lr1 = rmq_amd.gen_queries(1 << 18, -6, n, 5)
    lr2 = rmq_amd.gen_queries(1 << 18, -1, n, 5)
    ref1 = oracle.CpuHierarchy(A, 64, 5, 16).query(lr1)[1]
    ref2 = oracle.CpuHierarchy(A, 64, 5, 16).query(lr2)[1]
    dA = torch.from_numpy(A).cuda()
    s1, s2 = torch.cuda.Stream(), torch.cuda.Stream()
    with rmq_amd.Hierarchy(dA, 64, 16) as h:
        torch.cuda.synchronize()
        q1 = torch.from_numpy(lr1.view(np.int32)).cuda()
        q2 = torch.from_numpy(lr2.view(np.int32)).cuda()
        torch.cuda.synchronize()
        _say("streams: concurrent queries")
        for _ in range(3):
            _, i1 = h.query(q1, False, True, stream=s1)
            _, i2 = h.query(q2, False, True, stream=s2)
        torch.cuda.synchronize()
        assert np.array_equal(i1.cpu().numpy().view(np.uint32), ref1)
        assert np.array_equal(i2.cpu().numpy().view(np.uint32), ref2)


def test_error_paths_on_device(device):
    import torch

    dA = torch.rand(1000, device="cuda")
    with pytest.raises(rmq_amd.RMQError) as e:
        rmq_amd.Hierarchy(dA, 63, 16)
    assert e.value.status == rmq_amd.RMQ_ETHRESHOLD
    with rmq_amd.Hierarchy(dA, 64, 16) as h:
        empty = torch.empty((0, 2), dtype=torch.int32, device="cuda")
        v, i = h.query(empty, True, True)
        assert v.numel() == 0 and i.numel() == 0
        L = rmq_amd.lib()
        q = torch.zeros((4, 2), dtype=torch.int32, device="cuda")
        assert L.rmq_query(h._h, q.data_ptr(), 4, None, None, None) == rmq_amd.RMQ_EINVAL


def test_check_queries_mode(device):
    """RMQ_BUILD_CHECK_QUERIES: valid queries are answered as without it; a set holding one query
    outside 0 <= l <= r < n is refused with RMQ_EINVAL and nothing is written (the bad query is
    never handed to a query kernel). Covers rmq_query and rmq_query64 (one piece and chunked)."""
    import torch

    rng = np.random.default_rng(17)
    n = 50000
    A = rng.integers(0, 7, size=n).astype(np.float32)
    lr = np.sort(rng.integers(0, n, size=(3000, 2)), axis=1).astype(np.uint32)
    ref_i = oracle.cpu_base_index(A, lr)
    dA = torch.from_numpy(A).cuda()
    for alg, opts in ((20, dict(cg_amount_log=0, pack=True)), (16, {}), (24, dict(group_lanes=4))):
        _say(f"check_queries alg={alg} opts={opts}")
        with rmq_amd.Hierarchy(dA, _min_t(alg, opts), alg, check_queries=True, **opts) as h:
            v, i = h.query(torch.from_numpy(lr.view(np.int32)).cuda(), True, True)
            torch.cuda.synchronize()
            assert_exact(A, lr, v.cpu().numpy(), i.cpu().numpy().view(np.uint32), ref_i, f"checked alg={alg}")
            for bad in ([5, 4], [0, n], [n, n], [n - 1, 0xFFFFFFFF]):
                lb = lr.copy()
                lb[1234] = bad
                out_v = torch.full((lb.shape[0],), 7.0, device="cuda")
                with pytest.raises(rmq_amd.RMQError) as e:
                    h.query(torch.from_numpy(lb.view(np.int32)).cuda(), True, False, out_val=out_v)
                assert e.value.status == rmq_amd.RMQ_EINVAL, bad
                assert bool((out_v == 7.0).all()), "a refused call wrote answers"
    lr64 = lr.astype(np.uint64)
    for clog in (0, 12):
        _say(f"check_queries query64 chunk_log={clog}")
        with rmq_amd.Hierarchy(dA, 16, 20, cg_amount_log=0, chunk_log=clog, check_queries=True) as h:
            v, i = h.query64(torch.from_numpy(lr64.view(np.int64)).cuda(), True, True)
            torch.cuda.synchronize()
            assert np.array_equal(i.cpu().numpy().view(np.uint64), ref_i.astype(np.uint64))
            lb = lr64.copy()
            lb[77] = [0, n]
            with pytest.raises(rmq_amd.RMQError) as e:
                h.query64(torch.from_numpy(lb.view(np.int64)).cuda(), True, True)
            assert e.value.status == rmq_amd.RMQ_EINVAL


@pytest.mark.timeout(400)
def test_uniform_lengths_2_26(device):
    """The north-star workload: n = q = 2^26, uniform lengths (lr = -1, the paper's sweep), value
    and index mode. Alg 16 at the reference default t = 2^15, alg 20 at R = 8 and the bench's
    uniform-length configuration (alg 20, R = 4, t = 2^10), packed and not, one or two queries
    per lane (two at t = 2^10: the top level and its table in LDS): 2^22 queries
    against the oracle hierarchy (pinned to CPU BASE in test_oracle.py) plus a CPU BASE sample;
    the full 2^26 queries by size-independent properties (every configuration gives the same
    answer, the index lies in [l, r], the value is A[index])."""
    import torch

    _CACHE.clear()
    n = q = 1 << 26
    A = _array_2_26()
    lr = rmq_amd.gen_queries(q, -1, n, SEED)
    sub = lr[: 1 << 22]
    ref_v, ref_i = oracle.CpuHierarchy(A, 1 << 15, 5, 16).query(sub)
    sample = rmq_amd.gen_sample_indices(sub.shape[0], 384, SEED + 5)
    assert np.array_equal(oracle.cpu_base_index(A, sub[sample]), ref_i[sample])
    dA = torch.from_numpy(A).cuda()
    dQ = torch.from_numpy(lr.view(np.int32)).cuda()
    dS = dQ[: sub.shape[0]]
    lo = torch.from_numpy(lr[:, 0].astype(np.int64)).cuda()
    hi = torch.from_numpy(lr[:, 1].astype(np.int64)).cuda()
    first = None
    for alg, t, opts in ((16, 1 << 15, {}), (20, 1 << 13, dict(cg_amount_log=1)),
                         (20, 1 << 10, dict(cg_amount_log=0)), (20, 1 << 10, dict(cg_amount_log=0, pack=True)),
                         (24, 1 << 15, dict(group_lanes=4)),
                         # two per lane: top level and table in LDS (t = 2^10), in L2 (alg 16), packed
                         (20, 1 << 10, dict(cg_amount_log=0, paired=True)), (16, 1 << 15, dict(paired=True)),
                         (20, 1 << 10, dict(cg_amount_log=0, paired=True, pack=True))):
        _say(f"uniform 2^26 alg={alg} t={t} opts={opts}")
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            v, i = h.query(dS, True, True)
            vv, _ = h.query(dS, True, False)
            torch.cuda.synchronize()
            assert_exact(A, sub, v.cpu().numpy(), i.cpu().numpy().view(np.uint32), ref_i,
                         f"uniform alg={alg} t={t} {opts}")
            assert_exact(A, sub, vv.cpu().numpy(), None, ref_i, f"uniform value alg={alg} t={t} {opts}")
            fv, fi = h.query(dQ, True, True)
            torch.cuda.synchronize()
        fi64 = fi.to(torch.int64) & 0xFFFFFFFF
        assert bool(((fi64 >= lo) & (fi64 <= hi)).all()), (alg, t, opts)
        assert torch.equal(fv.view(torch.int32), dA[fi64].view(torch.int32)), (alg, t, opts)
        if first is None:
            first = (fv, fi)
        else:
            assert torch.equal(fi, first[1]), (alg, t, opts)
            assert torch.equal(fv.view(torch.int32), first[0].view(torch.int32)), (alg, t, opts)
    del dA, dQ, lo, hi, first
    torch.cuda.empty_cache()


@pytest.mark.timeout(300)
def test_query_count_above_one_launch(device):
    """q = 2^28 + 4099 queries (more than one launch of the one-lane-per-query kernels holds):
    the call is split into launches of 2^28, and every answer equals GPU BASE's (itself split the
    same way), for alg 16 and the packed lane solver; a sample, the launch boundary and the last
    queries are checked against CPU BASE."""
    import torch

    _CACHE.clear()
    rng = np.random.default_rng(29)
    n = 100003
    A = rng.integers(0, 1000, size=n).astype(np.float32)  # many ties
    q = (1 << 28) + 4099
    dA = torch.from_numpy(A).cuda()
    g = torch.Generator(device="cuda")
    g.manual_seed(29)
    ln = torch.randint(1, 300, (q,), device="cuda", dtype=torch.int32, generator=g)
    l = (torch.rand(q, device="cuda", generator=g) * (n - ln).to(torch.float32)).to(torch.int32)
    l = torch.minimum(l, n - ln)
    dQ = torch.stack([l, l + ln - 1], dim=1).contiguous()
    del ln, l
    _say("q > 2^28: GPU BASE")
    _, bi = rmq_amd.gpu_base(dA, dQ, False, True)
    torch.cuda.synchronize()
    edge = np.concatenate([np.arange((1 << 28) - 3, (1 << 28) + 3), np.arange(q - 8, q),
                           rmq_amd.gen_sample_indices(q, 2000, 3).astype(np.int64)])
    lr_edge = dQ[torch.from_numpy(edge).cuda()].cpu().numpy().view(np.uint32)
    ref = oracle.cpu_base_index(A, lr_edge)
    assert np.array_equal(bi[torch.from_numpy(edge).cuda()].cpu().numpy().view(np.uint32), ref)
    for alg, t, opts in ((16, 64, {}), (20, 16, dict(cg_amount_log=0, pack=True))):
        _say(f"q > 2^28: alg={alg} t={t} opts={opts}")
        with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
            v, i = h.query(dQ, True, True)
            torch.cuda.synchronize()
        assert torch.equal(i, bi), (alg, opts)
        assert torch.equal(v.view(torch.int32), dA[i.to(torch.int64)].view(torch.int32)), (alg, opts)
        del v, i
    del dQ, bi
    torch.cuda.empty_cache()


def _nan_cases():
    import json
    import os

    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "nan_vectors.json")
    with open(path) as f:
        return json.load(f)


def test_nan_vectors(device):
    """Arrays holding NaN follow CPU BASE's rule (src/cpu_methods.h:7-16, 42-53): a NaN at A[l]
    answers l, anywhere else a NaN never wins. Every variant, a hierarchy of several levels and
    one of one level (the sparse table over the array itself), value + index, value only and
    index only, against the fixture's known answers (value bits name the position: every NaN
    has its own payload)."""
    for case in _nan_cases():
        A = np.array(case["array_bits"], dtype=np.uint32).view(np.float32)
        lr = np.array(case["lr"], dtype=np.uint32)
        want_i = np.array(case["index"], dtype=np.uint32)
        want_v = np.array(case["value_bits"], dtype=np.uint32)
        for alg, opts in VARIANTS:
            for t in (_min_t(alg, opts), 1 << 15):
                _say(f"nan {case['name']} alg={alg} opts={opts} t={t}")
                v, i = run_gpu(A, lr, alg, t, True, True, **opts)
                assert np.array_equal(i, want_i), (case["name"], alg, opts, t, np.nonzero(i != want_i)[0][:5])
                assert np.array_equal(v.view(np.uint32), want_v), (case["name"], alg, opts, t)
                v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
                assert np.array_equal(v.view(np.uint32), want_v), (case["name"], alg, opts, t, "value")
                _, i = run_gpu(A, lr, alg, t, False, True, **opts)
                assert np.array_equal(i, want_i), (case["name"], alg, opts, t, "index")


@pytest.mark.parametrize("pattern", ["scattered", "runs"])
def test_nan_random_against_oracle(device, pattern):
    """n = 2^20 + 3 with NaN planted (1% scattered, or runs of 1..5000 covering whole blocks of
    every level), mixed lengths plus queries that start, end or sit next to a NaN: every variant,
    value and index, against the oracle's CPU BASE; GPU BASE agrees too."""
    import torch

    rng = np.random.default_rng({"scattered": 1, "runs": 2}[pattern])
    n = (1 << 20) + 3
    A = rmq_amd.gen_array(n, 5)
    if pattern == "scattered":
        pos = rng.choice(n, n // 100, replace=False)
    else:
        pos = np.concatenate([np.arange(s, min(n, s + int(L))) for s, L in
                              zip(rng.integers(0, n, 300), np.exp2(rng.uniform(0, 12.3, 300)))])
    A[pos] = np.uint32(0x7FC00000 | 77).view(np.float32)
    near = pos[rng.choice(pos.size, 3000)].astype(np.int64)
    lens = np.exp2(rng.uniform(0, 20, near.size)).astype(np.int64)
    starts = np.clip(near - rng.integers(0, 2, near.size) * lens, 0, n - 1)
    ends = np.clip(starts + lens, 0, n - 1)
    lr = np.concatenate([np.stack([starts, ends], 1), np.stack([near, near], 1),
                         rmq_amd.gen_queries(1 << 13, -6, n, 9).astype(np.int64)]).astype(np.uint32)
    lr = np.sort(lr, axis=1)
    ref_i = oracle.cpu_base_index(A, lr)
    for alg, opts in VARIANTS:
        t = 1 << 13 if alg == 20 else 1 << 15
        _say(f"nan random {pattern} alg={alg} opts={opts}")
        v, i = run_gpu(A, lr, alg, t, True, True, **opts)
        assert np.array_equal(i, ref_i), (alg, opts, np.nonzero(i != ref_i)[0][:5])
        assert np.array_equal(v.view(np.uint32), A[ref_i].view(np.uint32)), (alg, opts)
        v, _ = run_gpu(A, lr, alg, t, True, False, **opts)
        assert np.array_equal(v.view(np.uint32), A[ref_i].view(np.uint32)), (alg, opts, "value")
    short = np.nonzero((lr[:, 1].astype(np.int64) - lr[:, 0]) < (1 << 16))[0]
    sq = torch.from_numpy(np.ascontiguousarray(lr[short]).view(np.int32)).cuda()
    _, bi = rmq_amd.gpu_base(torch.from_numpy(A).cuda(), sq, False, True)
    torch.cuda.synchronize()
    assert np.array_equal(bi.cpu().numpy().view(np.uint32), ref_i[short])


@pytest.mark.parametrize("kind", ["uniform", "ties", "zeros", "nan"])
def test_lds_top_paired_kernel(device, kind):
    """Two queries per lane at R = 4 with the top level and its sparse table copied into LDS and
    the descending queries parked per wavefront (RMQ_BUILD_PAIRED, tops of 16..1024 entries), value
    and index mode, against the oracle's CPU BASE: uniform and mixed lengths, an odd query count,
    tops of 16, 256 and 1024 entries (n = 2^18) or 65 and 257 (n = 2^18 + 3, n % 4 != 0); values
    drawn from {0, 1, 2, 3} (guards equal to the best open most tails, so the park queues fill
    and are walked many times); +0 and -0 among 1 and 2 (value mode must return the leftmost zero's
    sign); NaN planted at 0.5% of the positions."""
    rng = np.random.default_rng({"uniform": 1, "ties": 2, "zeros": 4, "nan": 3}[kind])
    n = (1 << 18) + (3 if kind == "nan" else 0)
    A = rmq_amd.gen_array(n, 11)
    if kind == "ties":
        A = np.floor(A * 4).astype(np.float32)
    if kind == "zeros":
        A = np.array([0.0, -0.0, 1.0, 2.0], dtype=np.float32)[rng.integers(0, 4, n)]
    if kind == "nan":
        A[rng.choice(n, n // 200, replace=False)] = np.uint32(0x7FC00000 | 5).view(np.float32)
    lr = np.concatenate([rmq_amd.gen_queries((1 << 14) + 1, -1, n, 3), rmq_amd.gen_queries(1 << 12, -6, n, 4)])
    ref_i = oracle.cpu_base_index(A, lr)
    for t in (16, 256, 1024):
        _say(f"lds top {kind} t={t}")
        v, i = run_gpu(A, lr, 20, t, True, True, cg_amount_log=0, paired=True)
        assert_exact(A, lr, v, i, ref_i, f"lds top {kind} t={t}")
        v, _ = run_gpu(A, lr, 20, t, True, False, cg_amount_log=0, paired=True)
        assert_exact(A, lr, v, None, ref_i, f"lds top {kind} t={t} value")


@pytest.mark.parametrize("form", ["lane", "packed", "paired", "paired_packed", "lds_top", "alg16"])
def test_query_buffers_off_16_byte_alignment(device, form):
    """Query pairs that start 8 bytes past a 16-byte boundary (the two-per-lane kernels then load
    each pair on its own instead of both in one 16-byte load) and answer buffers 4 bytes past one,
    against the oracle's CPU BASE, value and index mode, with an odd query count."""
    import torch

    n = (1 << 18) + 1
    A = rmq_amd.gen_array(n, 5)
    lr = np.concatenate([rmq_amd.gen_queries(3001, -1, n, 6), rmq_amd.gen_queries(3000, -6, n, 7)])
    q = lr.shape[0]
    ref_i = oracle.cpu_base_index(A, lr)
    alg, t, opts = {"lane": (20, 1 << 13, dict(cg_amount_log=0)),
                    "packed": (20, 1 << 13, dict(cg_amount_log=0, pack=True)),
                    "paired": (20, 1 << 13, dict(cg_amount_log=1, paired=True)),
                    "paired_packed": (20, 1 << 13, dict(cg_amount_log=0, paired=True, pack=True)),
                    "lds_top": (20, 1 << 10, dict(cg_amount_log=0, paired=True)),
                    "alg16": (16, 1 << 15, dict(paired=True))}[form]
    dA = torch.from_numpy(A).cuda()
    big = torch.zeros((q + 1, 2), dtype=torch.int32, device="cuda")
    big[1:] = torch.from_numpy(lr.view(np.int32)).cuda()
    dQ = big[1:]
    ov = torch.full((q + 1,), -1.0, dtype=torch.float32, device="cuda")
    oi = torch.full((q + 1,), -1, dtype=torch.int32, device="cuda")
    assert dQ.data_ptr() % 16 == 8 and ov[1:].data_ptr() % 8 == 4 and oi[1:].data_ptr() % 8 == 4
    _say(f"unaligned buffers form={form}")
    with rmq_amd.Hierarchy(dA, t, alg, **opts) as h:
        h.query(dQ, True, True, ov[1:], oi[1:])
        torch.cuda.synchronize()
    assert ov[0].item() == -1.0 and oi[0].item() == -1  # nothing written before the buffers
    assert_exact(A, lr, ov[1:].cpu().numpy(), oi[1:].cpu().numpy().view(np.uint32), ref_i, f"unaligned {form}")
