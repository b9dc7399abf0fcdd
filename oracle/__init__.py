"""CPU oracle of the hierarchical RMQ path — TEST INFRASTRUCTURE ONLY.

Loaded only by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, and only as the
checker or the timed CPU baseline; the product (librmq.so, rmq_amd, rtxrmq) never imports it.
Binds oracle/liboracle.so (oracle/oracle.cpp, built by oracle/Makefile). Each function cites the
reference lines it restates in oracle.cpp.
"""

from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
MAX_LEVELS = 20

_lib = None
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_vp = ctypes.c_void_p


def build(mutant: int = 0, out_path: str = LIB_PATH) -> str:
    """Compile oracle.cpp; mutant > 0 plants a known mistake (tests only)."""
    if mutant == 0 and out_path == LIB_PATH:
        subprocess.run(["make", "-s", "-B", "-C", HERE], check=True)
    else:
        subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-fopenmp", "-shared",
                        f"-DORACLE_MUTANT={mutant}", "-o", out_path,
                        os.path.join(HERE, "oracle.cpp")], check=True)
    return out_path


def load(path: str) -> ctypes.CDLL:
    h = ctypes.CDLL(path)
    h.oracle_cpu_base_value.argtypes = [_vp, _u64, _vp, _u64, _vp, ctypes.c_int]
    h.oracle_cpu_base_value.restype = None
    h.oracle_cpu_base_index.argtypes = [_vp, _u64, _vp, _u64, _vp, ctypes.c_int]
    h.oracle_cpu_base_index.restype = None
    h.oracle_plan.argtypes = [_u64, _u64, _u32, _vp, _vp, _vp]
    h.oracle_plan.restype = _u64
    h.oracle_build.argtypes = [_vp, _u32, _u32, _vp, _vp, _vp, _vp]
    h.oracle_build.restype = ctypes.c_int
    h.oracle_hier_query.argtypes = [ctypes.c_int, _u32, _vp, _u32, _vp, _vp, _vp, _vp, _vp,
                                    _u64, _vp, _vp, ctypes.c_int]
    h.oracle_hier_query.restype = ctypes.c_int
    h.oracle_cover.argtypes = [ctypes.c_int, _u32, _u32, _vp, _vp, _u32, _u32, _vp]
    h.oracle_cover.restype = ctypes.c_int
    h.oracle_num_threads.argtypes = []
    h.oracle_num_threads.restype = ctypes.c_int
    return h


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


def num_threads() -> int:
    return int(lib().oracle_num_threads())


def _c(a, dtype):
    a = np.ascontiguousarray(a, dtype=dtype)
    return a, a.ctypes.data


def cpu_base_value(A: np.ndarray, lr: np.ndarray, nthreads: int = 0, L=None) -> np.ndarray:
    """alg 0: min value per query (src/cpu_methods.h:7-39)."""
    A, pa = _c(A, np.float32)
    lr, pq = _c(lr, np.uint32)
    q = lr.size // 2
    out = np.empty(q, dtype=np.float32)
    (L or lib()).oracle_cpu_base_value(pa, A.size, pq, q, out.ctypes.data, nthreads)
    return out


def cpu_base_index(A: np.ndarray, lr: np.ndarray, nthreads: int = 0, L=None) -> np.ndarray:
    """alg 100: leftmost argmin per query (src/cpu_methods.h:41-77)."""
    A, pa = _c(A, np.float32)
    lr, pq = _c(lr, np.uint32)
    q = lr.size // 2
    out = np.empty(q, dtype=np.uint32)
    (L or lib()).oracle_cpu_base_index(pa, A.size, pq, q, out.ctypes.data, nthreads)
    return out


def plan(n: int, t: int, R_log: int, L=None):
    """Level sizes/offsets and buffer length (src/xxx.h:225-243)."""
    nl = np.zeros(1, dtype=np.uint32)
    sizes = np.zeros(MAX_LEVELS, dtype=np.uint64)
    offs = np.zeros(MAX_LEVELS, dtype=np.uint64)
    floats = (L or lib()).oracle_plan(n, t, R_log, nl.ctypes.data, sizes.ctypes.data, offs.ctypes.data)
    k = int(nl[0])
    return k, sizes[:k].copy(), offs[:k].copy(), int(floats)


class CpuHierarchy:
    """The hierarchy built on the CPU (src/xxx.h:91-120) plus leftmost indices."""

    def __init__(self, A: np.ndarray, t: int, R_log: int, kind: int = 16, L=None):
        self.A = np.ascontiguousarray(A, dtype=np.float32)
        self.R_log = R_log
        self.kind = kind
        self.L = L or lib()
        self.num_levels, self.sizes, self.offsets, self.floats = plan(self.A.size, t, R_log, self.L)
        self.lv = np.zeros(max(self.floats, 1), dtype=np.float32)
        self.li = np.zeros(max(self.floats, 1), dtype=np.uint32)
        st = self.L.oracle_build(self.A.ctypes.data, R_log, self.num_levels, self.sizes.ctypes.data,
                                self.offsets.ctypes.data, self.lv.ctypes.data, self.li.ctypes.data)
        assert st == 0

    def level(self, k: int):
        if k == 0:
            return self.A, np.arange(self.A.size, dtype=np.uint32)
        o, s = int(self.offsets[k]), int(self.sizes[k])
        return self.lv[o:o + s], self.li[o:o + s]

    def query(self, lr: np.ndarray, nthreads: int = 0):
        lr, pq = _c(lr, np.uint32)
        q = lr.size // 2
        val = np.empty(q, dtype=np.float32)
        idx = np.empty(q, dtype=np.uint32)
        st = self.L.oracle_hier_query(self.kind, self.R_log, self.A.ctypes.data, self.num_levels,
                                     self.sizes.ctypes.data, self.offsets.ctypes.data,
                                     self.lv.ctypes.data, self.li.ctypes.data, pq, q,
                                     val.ctypes.data, idx.ctypes.data, nthreads)
        assert st == 0
        return val, idx

    def cover(self, l: int, r: int) -> np.ndarray:
        c = np.zeros(self.A.size, dtype=np.uint32)
        st = self.L.oracle_cover(self.kind, self.R_log, self.num_levels, self.sizes.ctypes.data,
                                self.offsets.ctypes.data, l, r, c.ctypes.data)
        assert st == 0
        return c
