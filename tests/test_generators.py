"""The driver's seeded input generators (include/rmq_gen.h) against committed fixtures and an
independent pure-Python restatement (tests/golden/pygen.py). CPU only."""

import hashlib
import json
import math
import os

import numpy as np
import pytest

import pygen
import rmq_amd

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def fixtures():
    with open(os.path.join(GOLDEN, "generator_fixtures.json")) as f:
        return json.load(f)


def test_mt19937_known_answer():
    # C++ standard [rand.predef]: 10000th output of default-constructed mt19937 (seed 5489).
    g = pygen.MT(5489)
    for _ in range(9999):
        g()
    assert g() == 4123659995


def test_lognormal_scales_match_libm():
    # (uint32)pow(n, 0.6 / 0.3) as evaluated by glibc, src/rand.cuh:200,202
    assert int(math.pow(float(1 << 26), 0.6)) == 49667
    assert int(math.pow(float(1 << 30), 0.3)) == 511
    assert int(math.pow(float(1 << 30), 0.6)) == 262143
    assert int(math.pow(float(1 << 20), 0.6)) == 4095
    assert int(math.pow(float(1 << 20), 0.3)) == 63


def test_array_fixtures(fixtures):
    for case in fixtures["arrays"]:
        a = rmq_amd.gen_array(case["n"], case["seed"])
        assert a[:16].view(np.uint32).tolist() == case["head_bits"]
        assert _sha(a) == case["sha256"], case


def test_query_fixtures(fixtures):
    for case in fixtures["queries"]:
        out = rmq_amd.gen_queries(case["q"], case["lr"], case["n"], case["seed"],
                                  case["trivial_check"])
        assert out[:16].tolist() == case["head"], case
        assert _sha(out) == case["sha256"], case


def test_sample_fixtures(fixtures):
    for case in fixtures["samples"]:
        s = rmq_amd.gen_sample_indices(case["original_size"], case["sample_size"], case["seed"])
        assert s[:16].tolist() == case["head"]
        assert _sha(s) == case["sha256"]


@pytest.mark.parametrize("lr", [-1, -2, -3, -4, -5, -6, 1, 7, 64])
def test_random_tuples_against_restatement(lr):
    rng = np.random.default_rng(abs(lr) * 1000 + 3)
    for _ in range(4):
        n = int(rng.integers(max(lr, 1), 1 << 24))
        seed = int(rng.integers(0, 1 << 31))
        q = 600
        a = rmq_amd.gen_queries(q, lr, n, seed)
        b = pygen.gen_queries(q, lr, n, seed)
        assert np.array_equal(a, b), (lr, n, seed)


@pytest.mark.parametrize("lr", [-1, -2, -6, 7])
def test_queries64_equal_32bit_below_2e32(lr):
    """The IS_LONG generator draws through the same 64-bit unsigned type: identical pairs."""
    for n, seed in ((1000, 3), (1 << 24, 27722), ((1 << 32) - 5, 11)):
        a = rmq_amd.gen_queries(500, lr, n, seed).astype(np.uint64)
        b = rmq_amd.gen_queries64(500, lr, n, seed)
        assert np.array_equal(a, b), (lr, n)


@pytest.mark.parametrize("lr", [-1, -2, -3, -6, 7])
def test_queries64_above_2e32_against_restatement(lr):
    """n > 2^32: 64-bit draws combine two 32-bit outputs (the restatement's wide branch)."""
    for n, seed in (((1 << 32) + 17, 5), ((1 << 33) + 12345, 27722), ((1 << 37) + 3, 9)):
        a = rmq_amd.gen_queries64(300, lr, n, seed)
        b = pygen.gen_queries(300, lr, n, seed, dtype=np.uint64)
        assert np.array_equal(a, b), (lr, n)
        assert (a[:, 0] <= a[:, 1]).all() and (a[:, 1] < n).all()
    if lr == -1:  # uniform lengths reach far above 2^32 only through the wide branch
        assert int(rmq_amd.gen_queries64(300, -1, (1 << 37) + 3, 9)[:, 1].max()) > (1 << 33)


def test_query_invariants():
    for lr in (-1, -2, -3, -6, 5):
        n = 100003
        Q = rmq_amd.gen_queries(20000, lr, n, 27722)
        assert (Q[:, 0] <= Q[:, 1]).all() and (Q[:, 1] < n).all()
    # the seed quirk: only lr = -6 depends on the seed (src/rand.cuh:70,91,112,131)
    assert np.array_equal(rmq_amd.gen_queries(100, -2, 1 << 20, 1),
                          rmq_amd.gen_queries(100, -2, 1 << 20, 99))
    assert not np.array_equal(rmq_amd.gen_queries(100, -6, 1 << 20, 1),
                              rmq_amd.gen_queries(100, -6, 1 << 20, 99))


def test_invalid_generator_args():
    with pytest.raises(ValueError):
        rmq_amd.gen_queries(10, 0, 100, 1)
    with pytest.raises(ValueError):
        rmq_amd.gen_queries(10, 101, 100, 1)
    with pytest.raises(ValueError):
        rmq_amd.gen_queries(10, -7, 100, 1)
