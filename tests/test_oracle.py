"""Pins the CPU oracle (oracle/oracle.cpp) before anything is checked against it. CPU only.

Each check compares the oracle with something independent of it: numpy's min/argmin (argmin
returns the first occurrence, i.e. the leftmost rule), hand-built tie/±0/inf vectors, literal
metadata known answers, a numpy block-minimum over each entry's level-0 span, and exact
coverage of [l, r] by the traversal. The planted-mistake tests build oracle variants with one
known error each (ORACLE_MUTANT) and require the matching check to fail on them.
"""

import json
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _tie_vectors():
    with open(os.path.join(GOLDEN, "tie_vectors.json")) as f:
        return json.load(f)


def _kats():
    with open(os.path.join(GOLDEN, "metadata_kats.json")) as f:
        return json.load(f)


def _random_case(rng, n, q, distinct=None):
    if distinct:
        A = rng.integers(0, distinct, size=n).astype(np.float32)
    else:
        A = rng.random(n, dtype=np.float32)
    lr = np.sort(rng.integers(0, n, size=(q, 2)), axis=1).astype(np.uint32)
    return A, lr


def _numpy_answers(A, lr):
    idx = np.array([l + int(np.argmin(A[l:r + 1])) for l, r in lr], dtype=np.uint32)
    return idx, A[idx]


def _nan_vectors():
    with open(os.path.join(GOLDEN, "nan_vectors.json")) as f:
        return json.load(f)


# ---- checks (return True when the oracle L passes) ------------------------------------------


def check_nan_vectors(L) -> bool:
    """CPU BASE on arrays holding NaN: a NaN at A[l] answers l, elsewhere it never wins (the
    fixture states the rule directly, tests/golden/make_nan_vectors.py)."""
    for case in _nan_vectors():
        A = np.array(case["array_bits"], dtype=np.uint32).view(np.float32)
        lr = np.array(case["lr"], dtype=np.uint32)
        if oracle.cpu_base_index(A, lr, L=L).tolist() != case["index"]:
            return False
        if oracle.cpu_base_value(A, lr, L=L).view(np.uint32).tolist() != case["value_bits"]:
            return False
    return True


def check_cpu_base_vs_numpy(L) -> bool:
    rng = np.random.default_rng(7)
    for n, distinct in ((1000, None), (1500, 4), (4099, 2), (64, None)):
        A, lr = _random_case(rng, n, 400, distinct)
        idx, val = _numpy_answers(A, lr)
        if not np.array_equal(oracle.cpu_base_index(A, lr, L=L), idx):
            return False
        got = oracle.cpu_base_value(A, lr, L=L)
        if not np.array_equal(got.view(np.uint32), val.view(np.uint32)):
            return False
    return True


def check_tie_vectors(L) -> bool:
    for case in _tie_vectors():
        A = np.array(case["array_bits"], dtype=np.uint32).view(np.float32)
        lr = np.array(case["lr"], dtype=np.uint32)
        if oracle.cpu_base_index(A, lr, L=L).tolist() != case["index"]:
            return False
        if oracle.cpu_base_value(A, lr, L=L).view(np.uint32).tolist() != case["value_bits"]:
            return False
        for kind, R_log, t in ((16, 5, 64), (20, 4, 16), (16, 3, 16), (24, 2, 8)):
            h = oracle.CpuHierarchy(A, t, R_log, kind, L=L)
            val, idx = h.query(lr)
            if idx.tolist() != case["index"] or val.view(np.uint32).tolist() != case["value_bits"]:
                return False
    return True


def check_metadata_kats(L) -> bool:
    for k in _kats():
        nl, sizes, offs, floats = oracle.plan(k["n"], k["t"], k["R_log"], L=L)
        if sizes.tolist() != k["sizes"] or offs.tolist() != k["offsets"] or floats != k["floats"]:
            return False
    return True


def check_level_contents(L) -> bool:
    rng = np.random.default_rng(11)
    for n, R_log, t, distinct in ((10007, 5, 64, None), (4096, 4, 16, 3), (777, 3, 8, None),
                                  (3001, 2, 4, 5)):
        A = rng.integers(0, distinct, size=n).astype(np.float32) if distinct else \
            rng.random(n, dtype=np.float32)
        h = oracle.CpuHierarchy(A, t, R_log, L=L)
        R = 1 << R_log
        for k in range(1, h.num_levels):
            vals, idxs = h.level(k)
            span = R ** k
            for j in range(vals.size):
                lo, hi = j * span, min((j + 1) * span, n)
                a = int(np.argmin(A[lo:hi])) + lo
                if idxs[j] != a or vals[j] != A[a]:
                    return False
    return True


def check_coverage(L) -> bool:
    rng = np.random.default_rng(5)
    for kind, R_log, t, n in ((16, 5, 64, 5000), (16, 3, 16, 3000), (20, 4, 16, 5000),
                              (20, 2, 4, 1000), (24, 5, 1 << 15, 2000), (16, 2, 8, 700)):
        h = oracle.CpuHierarchy(np.zeros(n, dtype=np.float32), t, R_log, kind, L=L)
        pairs = [tuple(sorted(int(x) for x in rng.integers(0, n, 2))) for _ in range(150)]
        pairs += [(0, n - 1), (0, 0), (n - 1, n - 1), (1, n - 2)]
        for l, r in pairs:
            c = h.cover(l, r)
            inside = c[l:r + 1]
            if c[:l].any() or c[r + 1:].any() or not inside.all():
                return False
            if kind != 20 and not (inside == 1).all():  # algs 16/24 partition [l, r]
                return False
    return True


def check_hier_vs_bruteforce(L) -> bool:
    rng = np.random.default_rng(9)
    for kind, R_log, t, n, distinct in ((16, 5, 64, 20000, None), (20, 4, 16, 9000, 6),
                                        (24, 5, 1 << 12, 30000, None), (20, 3, 8, 5000, None),
                                        (16, 6, 128, 40000, 2)):
        A, lr = _random_case(rng, n, 500, distinct)
        idx, val = _numpy_answers(A, lr)
        h = oracle.CpuHierarchy(A, t, R_log, kind, L=L)
        hv, hi = h.query(lr)
        if not np.array_equal(hi, idx) or not np.array_equal(hv.view(np.uint32), val.view(np.uint32)):
            return False
    return True


CHECKS = {
    "cpu_base_vs_numpy": check_cpu_base_vs_numpy,
    "tie_vectors": check_tie_vectors,
    "metadata_kats": check_metadata_kats,
    "level_contents": check_level_contents,
    "coverage": check_coverage,
    "hier_vs_bruteforce": check_hier_vs_bruteforce,
    "nan_vectors": check_nan_vectors,
}


@pytest.mark.parametrize("name", sorted(CHECKS))
def test_oracle_passes(name):
    assert CHECKS[name](oracle.lib())


# mutant id -> the check that must catch it (see ORACLE_MUTANT in oracle/oracle.cpp)
MUTANTS = {
    1: "tie_vectors",          # rightmost tie in CPU BASE IDX
    2: "cpu_base_vs_numpy",    # CPU BASE drops the last element
    3: "metadata_kats",        # level offsets not 32-aligned
    4: "level_contents",       # build block one element short
    5: "coverage",             # right partial block dropped by the traversal
    6: "tie_vectors",          # last equal candidate wins in the hierarchy
    7: "nan_vectors",          # CPU BASE seeded with +inf instead of A[l]
}


@pytest.mark.parametrize("mutant", sorted(MUTANTS))
def test_planted_mistake_is_caught(mutant, tmp_path):
    path = oracle.build(mutant, str(tmp_path / f"liboracle_m{mutant}.so"))
    L = oracle.load(path)
    assert not CHECKS[MUTANTS[mutant]](L), f"mutant {mutant} slipped past {MUTANTS[mutant]}"


def test_hier_bigger_sample_against_cpu_base():
    """Reference generator inputs: oracle hierarchy == oracle CPU BASE on a mid-size case."""
    import rmq_amd

    n, q = 1 << 20, 1 << 12
    A = rmq_amd.gen_array(n, 27722)
    for lr_mode in (-1, -2, -3, -6):
        lr = rmq_amd.gen_queries(q, lr_mode, n, 27722)
        ref_i = oracle.cpu_base_index(A, lr)
        ref_v = oracle.cpu_base_value(A, lr)
        assert np.array_equal(A[ref_i].view(np.uint32), ref_v.view(np.uint32))
        for kind, R_log, t in ((16, 5, 1 << 15), (20, 4, 1 << 15), (24, 5, 64), (20, 3, 32)):
            hv, hi = oracle.CpuHierarchy(A, t, R_log, kind).query(lr)
            assert np.array_equal(hi, ref_i)
            assert np.array_equal(hv.view(np.uint32), ref_v.view(np.uint32))
