"""C-ABI contract of librmq.so on the CPU: it loads, exports every function the public headers
declare, plans exactly the reference metadata, and rejects bad arguments with status codes
before touching a device (no compute calls here)."""

import ctypes
import json
import os
import re

import pytest

import oracle
import rmq_amd
from rmq_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


def _declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for fn in os.listdir(inc):
        if fn.endswith(".h"):
            text = open(os.path.join(inc, fn)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(rmq_[a-z_0-9]+)\s*\(", text))
    return names


def test_library_loads_and_exports_every_declared_symbol():
    L = rmq_amd.lib()
    declared = _declared_functions()
    assert {"rmq_build", "rmq_query", "rmq_free"} <= declared
    for name in sorted(declared):
        assert hasattr(L, name), f"{name} declared in include/ but not exported"
    # and the Python binding gives every declared function a signature
    assert declared <= set(_lib.SIGNATURES)
    assert rmq_amd.version().startswith("rmq-mi355x")


def _plan_for(R_log, n, t):
    # alg 16 with literal cg settings realises any R = 2^(g+c); alg 20 covers R = 4 and 8 too
    if R_log >= 3:
        return rmq_amd.plan(n, t, 16, cg_size_log=R_log, cg_amount_log=0)
    return rmq_amd.plan(n, t, 20, cg_size_log=0, cg_amount_log=R_log - 2)


def test_plan_matches_metadata_known_answers():
    with open(os.path.join(GOLDEN, "metadata_kats.json")) as f:
        kats = json.load(f)
    checked = 0
    for k in kats:
        R = 1 << k["R_log"]
        if k["t"] >= 2 * R and k["R_log"] >= 3:
            p = rmq_amd.plan(k["n"], k["t"], 16, cg_size_log=k["R_log"], cg_amount_log=0)
        elif k["t"] >= R and 2 <= k["R_log"] <= 6:
            p = rmq_amd.plan(k["n"], k["t"], 20, cg_size_log=0, cg_amount_log=k["R_log"] - 2)
        else:  # below every algorithm's threshold: only the oracle plans it
            with pytest.raises(rmq_amd.RMQError):
                _plan_for(k["R_log"], k["n"], k["t"])
            continue
        checked += 1
        assert p.sizes == k["sizes"], k
        assert p.offsets == k["offsets"], k
        assert p.buffer_floats == k["floats"], k
        assert p.R == 1 << k["R_log"]
    assert checked >= 10


def test_plan_matches_oracle_on_random_shapes():
    import numpy as np

    rng = np.random.default_rng(3)
    for _ in range(300):
        n = int(rng.integers(1, 1 << 31))
        R_log = int(rng.integers(3, 8))
        t = int(rng.integers(2 << R_log, 1 << 20))
        p = rmq_amd.plan(n, t, 16, cg_size_log=R_log, cg_amount_log=0)
        nl, sizes, offs, floats = oracle.plan(n, t, R_log)
        assert p.sizes == sizes.tolist() and p.offsets == offs.tolist()
        assert p.buffer_floats == floats


def test_reference_defaults():
    # algs 16/24: R = 2^(3+2) = 32; alg 20: R = 2^(2+2) = 16 (CMakeLists.txt:61-67)
    assert rmq_amd.plan(1 << 20, 1 << 15, 16).R == 32
    assert rmq_amd.plan(1 << 20, 1 << 15, 24).R == 32
    assert rmq_amd.plan(1 << 20, 1 << 15, 20).R == 16
    # paper settings: g = 4, c = 1 -> R = 32; VL c = 1 -> R = 8
    assert rmq_amd.plan(1 << 20, 64, 16, cg_size_log=4, cg_amount_log=1).R == 32
    assert rmq_amd.plan(1 << 20, 32, 20, cg_amount_log=1).R == 8


@pytest.mark.parametrize("alg,R,ok_t", [(16, 32, 64), (24, 32, 64), (20, 16, 16)])
def test_threshold_rejection(alg, R, ok_t):
    # src/xxx.h:211-217 (t >= 2R) and src/xxx.h:938-944 (t >= R)
    with pytest.raises(rmq_amd.RMQError) as e:
        rmq_amd.plan(1 << 20, ok_t - 1, alg)
    assert e.value.status == rmq_amd.RMQ_ETHRESHOLD
    rmq_amd.plan(1 << 20, ok_t, alg)


def test_plan_of_the_64bit_form():
    """n >= 2^31 (the reference's IS_LONG build) plans with the same recurrence in 64 bits."""
    n = (1 << 34) + 5
    p = rmq_amd.plan(n, 1 << 15, 20, cg_amount_log=1)
    sizes = [n]
    while sizes[-1] > (1 << 15):
        sizes.append((sizes[-1] + 7) // 8)
    assert p.sizes == sizes and p.R == 8
    assert all(o % 32 == 0 for o in p.offsets[1:])


def test_execution_flags_leave_the_plan_unchanged():
    """RMQ_BUILD_PACK_DESCENTS and RMQ_BUILD_CHECK_QUERIES change how queries run, not the
    hierarchy: rmq_plan returns the same levels with and without them, for every algorithm."""
    L = _lib.lib()

    def plan(n, t, alg, c, flags):
        o = _lib.BuildOptions(3, c, 0, _lib.RMQ_BUILD_DEFAULT_CG | flags)
        nl, bf, R = ctypes.c_uint32(), ctypes.c_uint64(), ctypes.c_uint32()
        sizes = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
        offs = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
        st = L.rmq_plan(n, t, alg, ctypes.byref(o), ctypes.byref(nl), sizes, offs, ctypes.byref(bf),
                        ctypes.byref(R))
        assert st == rmq_amd.RMQ_OK
        return list(sizes[:nl.value]), list(offs[:nl.value]), bf.value, R.value

    every = _lib.RMQ_BUILD_PACK_DESCENTS | _lib.RMQ_BUILD_CHECK_QUERIES
    for n, t, alg, c in ((1 << 26, 1 << 13, 20, 0), (1 << 20, 1 << 15, 16, 2), (100003, 64, 24, 1),
                         ((1 << 33) + 3, 1 << 13, 20, 1)):
        base = plan(n, t, alg, c, 0)
        for flags in (_lib.RMQ_BUILD_PACK_DESCENTS, _lib.RMQ_BUILD_CHECK_QUERIES, every):
            assert plan(n, t, alg, c, flags) == base, (n, t, alg, c, flags)


def test_invalid_arguments():
    for n, t, alg in ((0, 64, 16), ((1 << 38) + 1, 64, 16), (100, 64, 17), (100, 0, 16), (100, 64, 2)):
        with pytest.raises(rmq_amd.RMQError) as e:
            rmq_amd.plan(n, t, alg)
        assert e.value.status == rmq_amd.RMQ_EINVAL
    with pytest.raises(rmq_amd.RMQError) as e:
        rmq_amd.plan(100, 64, 16, group_lanes=3)
    assert e.value.status == rmq_amd.RMQ_EUNSUPPORTED
    for g in (1, 4, 8, 16, 32, 64):  # a lane per query up to a whole wavefront per query
        rmq_amd.plan(100, 64, 16, group_lanes=g)
    # a lane per query runs the pruned order only, on R <= 64
    for kw in (dict(full=True), dict(cg_size_log=4, cg_amount_log=3)):
        with pytest.raises(rmq_amd.RMQError) as e:
            rmq_amd.plan(100, 256, 16, group_lanes=1, **kw)
        assert e.value.status == rmq_amd.RMQ_EUNSUPPORTED
    with pytest.raises(rmq_amd.RMQError) as e:
        rmq_amd.plan(100, 64, 16, group_lanes=128)
    assert e.value.status == rmq_amd.RMQ_EUNSUPPORTED


def test_null_and_misaligned_pointers_rejected_before_device_use():
    L = rmq_amd.lib()
    h = ctypes.c_void_p()
    assert L.rmq_build(None, 100, 64, 16, None, ctypes.byref(h)) == rmq_amd.RMQ_EINVAL
    assert h.value is None
    assert L.rmq_build(0x1004, 100, 64, 16, None, ctypes.byref(h)) == rmq_amd.RMQ_EINVAL
    assert L.rmq_build(0x1000, 100, 63, 16, None, ctypes.byref(h)) == rmq_amd.RMQ_ETHRESHOLD
    assert L.rmq_build(0x1000, 100, 64, 16, None, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_query(None, 0x1000, 10, 0x2000, None, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_gpu_base(None, 10, 0x1000, 10, 0x2000, None, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_gpu_base(0x1000, 10, 0x1004, 10, 0x2000, None, None) == rmq_amd.RMQ_EINVAL
    assert L.rmq_gpu_base(0x1000, 10, 0x1000, 0, None, None, None) == rmq_amd.RMQ_OK  # q = 0
    L.rmq_free(None)
    for s in (0, -1, -2, -3, -4, -5):
        assert rmq_amd.status_str(s)
