"""rmq_amd — Python host API over librmq.so (MI355X hierarchical range-minimum queries).

Device memory and streams come from PyTorch (plumbing only); every computation runs in the
library's HIP kernels through the C ABI of include/rmq.h. Host generators reproduce the
reference driver's seeded inputs (include/rmq_gen.h).

    import rmq_amd
    A = torch.from_numpy(rmq_amd.gen_array(n, seed)).cuda()
    Q = torch.from_numpy(rmq_amd.gen_queries(q, -2, n, seed)).cuda()
    with rmq_amd.Hierarchy(A, t=1 << 15, alg=16) as h:
        val, idx = h.query(Q, want_val=True, want_idx=True)
"""

from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _lib
from ._lib import (  # noqa: F401  (re-exported)
    RMQ_ALG_GPU_BASE,
    RMQ_ALG_VECTOR_LOAD,
    RMQ_ALG_XXX,
    RMQ_ALG_XXX_MULTI_LOAD,
    RMQ_EDEVICE,
    RMQ_EINVAL,
    RMQ_ENOMEM,
    RMQ_ETHRESHOLD,
    RMQ_EUNSUPPORTED,
    RMQ_EAGAIN,
    RMQ_OK,
    RMQError,
    BuildOptions,
    LIB_PATH,
    lib,
    status_str,
    version,
)

ALGS = (RMQ_ALG_XXX, RMQ_ALG_VECTOR_LOAD, RMQ_ALG_XXX_MULTI_LOAD)


@dataclass
class Plan:
    """Level metadata (the reference's xxx_metadata_type, src/constants.cuh:34-39)."""

    num_levels: int
    sizes: List[int]
    offsets: List[int]
    buffer_floats: int
    R: int


def _options(cg_size_log, cg_amount_log, group_lanes, full=False, chunk_log=0,
             top_table=True, pack=False, check=False, paired=False, scans=None) -> Optional[BuildOptions]:
    if (cg_size_log is None and cg_amount_log is None and not group_lanes and not full and not chunk_log
            and top_table and not pack and not check and not paired and scans is None):
        return None
    o = BuildOptions()
    o.cg_size_log = 3 if cg_size_log is None else int(cg_size_log)
    o.cg_amount_log = 2 if cg_amount_log is None else int(cg_amount_log)
    o.group_lanes = int(group_lanes or 0)
    o.flags = (_lib.RMQ_BUILD_DEFAULT_CG | (_lib.RMQ_BUILD_FULL if full else 0) |
               (0 if top_table else _lib.RMQ_BUILD_NO_TOP_TABLE) |
               (_lib.RMQ_BUILD_PACK_DESCENTS if pack else 0) |
               (_lib.RMQ_BUILD_CHECK_QUERIES if check else 0) | (_lib.RMQ_BUILD_PAIRED if paired else 0) |
               (_lib.RMQ_BUILD_BLOCK_SCANS if scans is True else 0) |
               (_lib.RMQ_BUILD_NO_BLOCK_SCANS if scans is False else 0) |
               _lib.RMQ_BUILD_CHUNK_LOG(chunk_log))
    return o


def _opt_ptr(o: Optional[BuildOptions]):
    return ctypes.byref(o) if o is not None else None


def plan(n: int, t: int, alg: int = RMQ_ALG_XXX, cg_size_log=None, cg_amount_log=None,
         group_lanes: int = 0, full: bool = False) -> Plan:
    """Host-only level plan (src/xxx.h:225-243). Raises RMQError on invalid input."""
    L = _lib.lib()
    nl = ctypes.c_uint32()
    sizes = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
    offs = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
    bf = ctypes.c_uint64()
    R = ctypes.c_uint32()
    o = _options(cg_size_log, cg_amount_log, group_lanes, full)
    _lib.check(L.rmq_plan(n, t, alg, _opt_ptr(o), ctypes.byref(nl), sizes, offs, ctypes.byref(bf),
                          ctypes.byref(R)), "rmq_plan")
    k = nl.value
    return Plan(k, list(sizes[:k]), list(offs[:k]), bf.value, R.value)


# ---- host generators (include/rmq_gen.h) ---------------------------------------------------

def gen_array(n: int, seed: int) -> np.ndarray:
    """float32[n] from std::mt19937(seed) + uniform_real_distribution<float>(0,1)."""
    out = np.empty(n, dtype=np.float32)
    _lib.lib().rmq_gen_array(n, seed, out.ctypes.data)
    return out


def gen_queries(q: int, lr: int, n: int, seed: int, trivial_check: bool = False) -> np.ndarray:
    """uint32[q, 2] inclusive (l, r) pairs exactly as the reference driver makes them."""
    out = np.empty((q, 2), dtype=np.uint32)
    st = _lib.lib().rmq_gen_queries(q, lr, n, seed, int(bool(trivial_check)), out.ctypes.data)
    if st != 0:
        raise ValueError(f"invalid query generator arguments: lr={lr}, n={n}")
    return out


def gen_queries64(q: int, lr: int, n: int, seed: int, trivial_check: bool = False) -> np.ndarray:
    """uint64[q, 2] pairs as the reference's IS_LONG build makes them (equal to gen_queries
    while n <= 2^32)."""
    out = np.empty((q, 2), dtype=np.uint64)
    st = _lib.lib().rmq_gen_queries64(q, lr, n, seed, int(bool(trivial_check)), out.ctypes.data)
    if st != 0:
        raise ValueError(f"invalid query generator arguments: lr={lr}, n={n}")
    return out


def gen_sample_indices(original_size: int, sample_size: int, seed: int) -> np.ndarray:
    out = np.empty(min(original_size, sample_size), dtype=np.uint32)
    k = _lib.lib().rmq_gen_sample_indices(original_size, sample_size, seed, out.ctypes.data)
    return out[:k]


# ---- device side -----------------------------------------------------------------------------

def _torch():
    import torch  # deferred: host-only users do not need it

    return torch


def _stream_handle(stream) -> Optional[int]:
    torch = _torch()
    if stream is None:
        stream = torch.cuda.current_stream()
    return stream.cuda_stream if hasattr(stream, "cuda_stream") else int(stream)


def _check_device_tensor(x, name: str, dtypes) -> None:
    torch = _torch()
    if not isinstance(x, torch.Tensor) or not x.is_cuda:
        raise ValueError(f"{name} must be a CUDA (HIP) tensor")
    if x.dtype not in dtypes:
        raise ValueError(f"{name} has dtype {x.dtype}, expected one of {dtypes}")
    if not x.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


class Hierarchy:
    """A built hierarchy resident in HBM (rmq_build / rmq_query / rmq_free).

    The array tensor is borrowed for the lifetime of the object, as the C ABI requires.
    Options beyond the reference's: group_lanes (lanes per query, algs 16/24), full (the
    reference's traversal order), chunk_log (the 64-bit form's chunk size), top_table=False (no
    sparse table over the top level), pack=True (RMQ_BUILD_PACK_DESCENTS), check_queries=True
    (RMQ_BUILD_CHECK_QUERIES: out-of-range queries raise RMQError instead of being answered),
    paired=True (RMQ_BUILD_PAIRED: two queries per lane), scans=True / False (RMQ_BUILD_BLOCK_SCANS /
    RMQ_BUILD_NO_BLOCK_SCANS: block prefix / suffix minima for the lane and group solvers; None = the
    library's default, built where R >= 32).
    """

    def __init__(self, array, t: int, alg: int = RMQ_ALG_XXX, cg_size_log=None,
                 cg_amount_log=None, group_lanes: int = 0, full: bool = False, chunk_log: int = 0,
                 top_table: bool = True, pack: bool = False, check_queries: bool = False, paired: bool = False,
                 scans: Optional[bool] = None, stream=None):
        torch = _torch()
        _check_device_tensor(array, "array", (torch.float32,))
        if array.dim() != 1:
            raise ValueError("array must be 1-D")
        self.array = array
        self.n = array.numel()
        self.t = int(t)
        self.alg = int(alg)
        self._h = ctypes.c_void_p()
        o = _options(cg_size_log, cg_amount_log, group_lanes, full, chunk_log, top_table, pack, check_queries,
                     paired, scans)
        st = _lib.lib().rmq_build_ex(array.data_ptr(), self.n, self.t, self.alg, _opt_ptr(o),
                                     _stream_handle(stream), ctypes.byref(self._h))
        _lib.check(st, "rmq_build")

    def query(self, lr, want_val: bool = True, want_idx: bool = False, out_val=None, out_idx=None,
              stream=None) -> Tuple[Optional["object"], Optional["object"]]:
        """Answer the queries lr (int32/uint32 tensor [q, 2], inclusive l, r) on the device.

        Returns (values float32[q] or None, leftmost indices int32[q] or None). Indices are
        stored as int32 tensors holding the uint32 values (n < 2^31).
        """
        torch = _torch()
        if self._h.value is None:
            raise RuntimeError("hierarchy already freed")
        _check_device_tensor(lr, "lr", (torch.int32, torch.uint32))
        q = lr.numel() // 2
        if lr.numel() != 2 * q:
            raise ValueError("lr must hold q pairs")
        if want_val and out_val is None:
            out_val = torch.empty(q, dtype=torch.float32, device=lr.device)
        if want_idx and out_idx is None:
            out_idx = torch.empty(q, dtype=torch.int32, device=lr.device)
        vp = out_val.data_ptr() if want_val else None
        ip = out_idx.data_ptr() if want_idx else None
        st = _lib.lib().rmq_query(self._h, lr.data_ptr(), q, vp, ip, _stream_handle(stream))
        _lib.check(st, "rmq_query")
        return (out_val if want_val else None, out_idx if want_idx else None)

    def query64(self, lr, want_val: bool = True, want_idx: bool = False, stream=None):
        """64-bit queries (int64/uint64 tensor [q, 2]) through rmq_query64: (values float32[q],
        leftmost indices int64[q]) — the only query of a hierarchy built in chunks."""
        torch = _torch()
        if self._h.value is None:
            raise RuntimeError("hierarchy already freed")
        _check_device_tensor(lr, "lr", (torch.int64, torch.uint64))
        q = lr.numel() // 2
        val = torch.empty(q, dtype=torch.float32, device=lr.device) if want_val else None
        idx = torch.empty(q, dtype=torch.int64, device=lr.device) if want_idx else None
        st = _lib.lib().rmq_query64(self._h, lr.data_ptr(), q, val.data_ptr() if want_val else None,
                                    idx.data_ptr() if want_idx else None, _stream_handle(stream))
        _lib.check(st, "rmq_query64")
        return val, idx

    def metadata(self) -> Plan:
        nl = ctypes.c_uint32()
        sizes = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
        offs = (ctypes.c_uint64 * _lib.RMQ_MAX_LEVELS)()
        bf = ctypes.c_uint64()
        R = ctypes.c_uint32()
        _lib.check(_lib.lib().rmq_get_metadata(self._h, ctypes.byref(nl), sizes, offs,
                                               ctypes.byref(bf), ctypes.byref(R)),
                   "rmq_get_metadata")
        k = nl.value
        return Plan(k, list(sizes[:k]), list(offs[:k]), bf.value, R.value)

    def levels(self, stream=None):
        """Levels 1.. read back (rmq_read_levels): (values float32[buffer_floats], leftmost
        level-0 indices int32[buffer_floats]), entry j of level k at metadata().offsets[k] + j.
        Alignment gaps between levels hold zeros here."""
        torch = _torch()
        if self._h.value is None:
            raise RuntimeError("hierarchy already freed")
        bf = self.metadata().buffer_floats
        val = torch.zeros(max(bf, 1), dtype=torch.float32, device=self.array.device)
        idx = torch.zeros(max(bf, 1), dtype=torch.int32, device=self.array.device)
        if bf:
            _lib.check(_lib.lib().rmq_read_levels(self._h, val.data_ptr(), idx.data_ptr(),
                                                  _stream_handle(stream)), "rmq_read_levels")
        return val[:bf], idx[:bf]

    def close(self) -> None:
        if self._h.value is not None:
            _lib.lib().rmq_free(self._h)
            self._h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PlainMany:
    """Per-array plain hierarchies built and queried from C loops (rmq_build_each /
    rmq_query_each): the baseline of the batched form. Offsets/sizes index one device buffer."""

    def __init__(self, arrays, offsets, sizes, t, alg: int = RMQ_ALG_XXX, stream=None):
        torch = _torch()
        _check_device_tensor(arrays, "arrays", (torch.float32,))
        self.count = count = len(sizes)
        # pointer and size tables as numpy arrays (no per-array Python work on the host path)
        ptrs = np.uint64(arrays.data_ptr()) + 4 * np.asarray(offsets, dtype=np.uint64)
        ns = np.asarray(sizes, dtype=np.uint64)
        ts = np.asarray(t, dtype=np.uint64)
        self._hs = (ctypes.c_void_p * count)()
        self.arrays = arrays
        _lib.check(_lib.lib().rmq_build_each(ptrs.ctypes.data, ns.ctypes.data, ts.ctypes.data, alg, None,
                                             count, _stream_handle(stream), self._hs), "rmq_build_each")

    def query(self, lr, qoffsets, out_idx, stream=None) -> None:
        """Index answers of every array's queries (local pairs in lr, array a's pairs at
        qoffsets[a]..qoffsets[a+1]) into out_idx at the same positions."""
        c = self.count
        qo = np.asarray(qoffsets, dtype=np.uint64)
        qs = np.uint64(lr.data_ptr()) + 8 * qo[:c]
        os_ = np.uint64(out_idx.data_ptr()) + 4 * qo[:c]
        qn = qo[1:c + 1] - qo[:c]
        _lib.check(_lib.lib().rmq_query_each(self._hs, qs.ctypes.data, qn.ctypes.data, None, os_.ctypes.data,
                                             c, _stream_handle(stream)), "rmq_query_each")

    def close(self) -> None:
        if getattr(self, "_hs", None) is not None:
            for h in self._hs:
                if h:
                    _lib.lib().rmq_free(h)
            self._hs = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_base(array, lr, want_val: bool = True, want_idx: bool = False, stream=None):
    """GPU BASE (algs 2 / 102, src/cuda_methods.cuh:4-68) through rmq_gpu_base."""
    torch = _torch()
    _check_device_tensor(array, "array", (torch.float32,))
    _check_device_tensor(lr, "lr", (torch.int32, torch.uint32))
    q = lr.numel() // 2
    val = torch.empty(q, dtype=torch.float32, device=lr.device) if want_val else None
    idx = torch.empty(q, dtype=torch.int32, device=lr.device) if want_idx else None
    st = _lib.lib().rmq_gpu_base(array.data_ptr(), array.numel(), lr.data_ptr(), q,
                                 val.data_ptr() if want_val else None,
                                 idx.data_ptr() if want_idx else None, _stream_handle(stream))
    _lib.check(st, "rmq_gpu_base")
    return val, idx


def gpu_base64(array, lr, want_val: bool = True, want_idx: bool = False, stream=None):
    """GPU BASE with 64-bit queries and indices (rmq_gpu_base64)."""
    torch = _torch()
    _check_device_tensor(array, "array", (torch.float32,))
    _check_device_tensor(lr, "lr", (torch.int64, torch.uint64))
    q = lr.numel() // 2
    val = torch.empty(q, dtype=torch.float32, device=lr.device) if want_val else None
    idx = torch.empty(q, dtype=torch.int64, device=lr.device) if want_idx else None
    st = _lib.lib().rmq_gpu_base64(array.data_ptr(), array.numel(), lr.data_ptr(), q,
                                   val.data_ptr() if want_val else None,
                                   idx.data_ptr() if want_idx else None, _stream_handle(stream))
    _lib.check(st, "rmq_gpu_base64")
    return val, idx


# ---- batched form (include/rmq_batch.h) ------------------------------------------------------

VARIANTS = {"lane": _lib.RMQ_VARIANT_LANE, "lane8": _lib.RMQ_VARIANT_LANE8, "lane4": _lib.RMQ_VARIANT_LANE4,
            "group4": _lib.RMQ_VARIANT_GROUP4, "group8": _lib.RMQ_VARIANT_GROUP8,
            "group16": _lib.RMQ_VARIANT_GROUP16, "group32": _lib.RMQ_VARIANT_GROUP32,
            "wave": _lib.RMQ_VARIANT_WAVE}


def fill_uniform(out, seed: int, start: int = 0, stream=None) -> None:
    """Fill a float32 CUDA tensor with the library's counter-based uniform [0, 1) numbers."""
    torch = _torch()
    _check_device_tensor(out, "out", (torch.float32,))
    _lib.check(_lib.lib().rmq_fill_uniform(out.data_ptr(), out.numel(), seed, start,
                                           _stream_handle(stream)), "rmq_fill_uniform")


# numpy layout of rmq_batch_array (include/rmq_batch.h), checked against the ctypes record
_BATCH_ARRAY_DTYPE = np.dtype([("offset", np.uint64), ("n", np.uint32), ("t", np.uint32),
                               ("variant", np.uint32), ("reserved", np.uint32)])
assert _BATCH_ARRAY_DTYPE.itemsize == ctypes.sizeof(_lib.BatchArray)


class Batch:
    """Many independent arrays, one build launch per level, one persistent query launch.

    arrays: float32 CUDA tensor holding every array; offsets/sizes (element units, offsets
    multiples of 4); t and variant per array (ints / names in VARIANTS).
    """

    def __init__(self, arrays, offsets, sizes, t, variant, stream=None):
        torch = _torch()
        _check_device_tensor(arrays, "arrays", (torch.float32,))
        count = len(sizes)
        # rmq_batch_array records, filled column-wise (a Python loop over ctypes fields costs
        # milliseconds at thousands of arrays)
        descs = np.zeros(count, dtype=_BATCH_ARRAY_DTYPE)
        descs["offset"] = np.asarray(offsets, dtype=np.uint64)
        descs["n"] = np.asarray(sizes, dtype=np.uint32)
        descs["t"] = np.asarray(t, dtype=np.uint32)
        descs["variant"] = [VARIANTS[v] if isinstance(v, str) else int(v) for v in variant]
        self.arrays = arrays
        self.count = count
        self._b = ctypes.c_void_p()
        self._plans = []
        _lib.check(_lib.lib().rmq_batch_build(arrays.data_ptr(), descs.ctypes.data, count,
                                              _stream_handle(stream), ctypes.byref(self._b)),
                   "rmq_batch_build")

    def plan(self, qoffsets, stream=None) -> "BatchPlan":
        q = np.ascontiguousarray(qoffsets, dtype=np.uint64)
        if q.size != self.count + 1:
            raise ValueError("qoffsets needs count + 1 entries")
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().rmq_batch_plan_queries(self._b, q.ctypes.data, _stream_handle(stream),
                                                     ctypes.byref(p)), "rmq_batch_plan_queries")
        plan = BatchPlan(self, p, int(q[-1]))
        self._plans.append(plan)
        return plan

    def level_bytes(self) -> int:
        return int(_lib.lib().rmq_batch_level_bytes(self._b))

    def levels(self, array: int, buffer_floats: int, stream=None):
        """Array `array`'s levels read back (rmq_batch_read_levels) into new tensors of
        buffer_floats entries (its rmq_plan buffer length): (values float32, indices int32)."""
        torch = _torch()
        dev = self.arrays.device
        val = torch.zeros(max(buffer_floats, 1), dtype=torch.float32, device=dev)
        idx = torch.zeros(max(buffer_floats, 1), dtype=torch.int32, device=dev)
        if buffer_floats:
            _lib.check(_lib.lib().rmq_batch_read_levels(self._b, array, val.data_ptr(), idx.data_ptr(),
                                                        _stream_handle(stream)), "rmq_batch_read_levels")
        return val[:buffer_floats], idx[:buffer_floats]

    def close(self) -> None:
        for p in self._plans:
            p.close()
        self._plans = []
        if self._b.value is not None:
            _lib.lib().rmq_batch_free(self._b)
            self._b = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class BatchPlan:
    """Device work queue for one query layout of a Batch."""

    def __init__(self, batch: Batch, handle, total: int):
        self.batch, self._p, self.total = batch, handle, total
        self._graphs = []

    def query(self, lr, want_val=True, want_idx=False, out_val=None, out_idx=None, stream=None):
        torch = _torch()
        _check_device_tensor(lr, "lr", (torch.int32, torch.uint32))
        if lr.numel() != 2 * self.total:
            raise ValueError("lr must hold the planned number of pairs")
        if want_val and out_val is None:
            out_val = torch.empty(self.total, dtype=torch.float32, device=lr.device)
        if want_idx and out_idx is None:
            out_idx = torch.empty(self.total, dtype=torch.int32, device=lr.device)
        _lib.check(_lib.lib().rmq_batch_query(self.batch._b, self._p, lr.data_ptr(),
                                              out_val.data_ptr() if want_val else None,
                                              out_idx.data_ptr() if want_idx else None,
                                              _stream_handle(stream)), "rmq_batch_query")
        return (out_val if want_val else None, out_idx if want_idx else None)

    def capture(self, lr, out_val=None, out_idx=None) -> "Graph":
        """Capture the query launch (counter reset + persistent kernel) into a HIP graph."""
        g = ctypes.c_void_p()
        _lib.check(_lib.lib().rmq_batch_capture(self.batch._b, self._p, lr.data_ptr(),
                                                out_val.data_ptr() if out_val is not None else None,
                                                out_idx.data_ptr() if out_idx is not None else None,
                                                None, ctypes.byref(g)), "rmq_batch_capture")
        graph = Graph(g, (lr, out_val, out_idx))
        self._graphs.append(graph)
        return graph

    def close(self) -> None:
        for g in self._graphs:
            g.close()
        self._graphs = []
        if self._p.value is not None:
            _lib.lib().rmq_batch_plan_free(self._p)
            self._p = ctypes.c_void_p()


class Graph:
    """An instantiated HIP graph of a batched query launch (buffers bound at capture)."""

    def __init__(self, handle, keep):
        self._g, self._keep = handle, keep

    def launch(self, stream=None) -> None:
        _lib.check(_lib.lib().rmq_graph_launch(self._g, _stream_handle(stream)), "rmq_graph_launch")

    def close(self) -> None:
        if self._g.value is not None:
            _lib.lib().rmq_graph_free(self._g)
            self._g = ctypes.c_void_p()


# ---- serving loop (include/rmq_serve.h) --------------------------------------------------------

class Server:
    """Resident hierarchies + continuous batching (single-threaded, driven by poll/drain).

    Requests keep their numpy buffers alive until completion: answers land in them.
    """

    def __init__(self, slots=0, max_batch_queries=0, reserve_bytes=0, max_arrays=0,
                 max_batch_requests=0, ring_queries=0):
        o = _lib.ServerOptions(slots, max_batch_queries, reserve_bytes, max_arrays,
                               max_batch_requests, ring_queries)
        self._s = ctypes.c_void_p()
        _lib.check(_lib.lib().rmq_server_create(ctypes.byref(o), ctypes.byref(self._s)),
                   "rmq_server_create")
        self._live = {}  # ticket -> (lr, val, idx) kept alive until done
        self._done_buf = np.zeros(1 << 16, dtype=np.uint64)

    def admit_uniform(self, n: int, seed: int, t: int = 64, variant="lane") -> int:
        return self._admit(_lib.RMQ_ADMIT_UNIFORM, None, n, seed, t, variant)

    def admit_host(self, array: np.ndarray, t: int = 64, variant="lane") -> int:
        a = np.ascontiguousarray(array, dtype=np.float32)
        return self._admit(_lib.RMQ_ADMIT_HOST, a.ctypes.data, a.size, 0, t, variant)

    def admit_device(self, array, t: int = 64, variant="lane") -> int:
        torch = _torch()
        _check_device_tensor(array, "array", (torch.float32,))
        return self._admit(_lib.RMQ_ADMIT_DEVICE, array.data_ptr(), array.numel(), 0, t, variant)

    def _admit(self, source, ptr, n, seed, t, variant) -> int:
        v = VARIANTS[variant] if isinstance(variant, str) else int(variant)
        aid = ctypes.c_int32()
        _lib.check(_lib.lib().rmq_server_admit(self._s, source, ptr, n, seed, t, v, ctypes.byref(aid)),
                   "rmq_server_admit")
        return aid.value

    def evict(self, array_id: int) -> None:
        _lib.check(_lib.lib().rmq_server_evict(self._s, array_id), "rmq_server_evict")

    def submit(self, array_id: int, lr: np.ndarray, want_val=True, want_idx=False,
               out_val: np.ndarray = None, out_idx: np.ndarray = None):
        """Queue queries on a resident array; returns (ticket, val, idx). The answers land in
        out_val / out_idx when given (contiguous float32 / uint32 of q entries, e.g. views of a
        reused pool), else in new arrays."""
        lr = np.ascontiguousarray(lr, dtype=np.uint32)
        q = lr.size // 2
        val = idx = None
        if want_val:
            val = np.empty(q, dtype=np.float32) if out_val is None else out_val
            if val.dtype != np.float32 or val.shape != (q,) or not val.flags.c_contiguous:
                raise ValueError("out_val must be a contiguous float32 array of q entries")
        if want_idx:
            idx = np.empty(q, dtype=np.uint32) if out_idx is None else out_idx
            if idx.dtype != np.uint32 or idx.shape != (q,) or not idx.flags.c_contiguous:
                raise ValueError("out_idx must be a contiguous uint32 array of q entries")
        t = ctypes.c_uint64()
        _lib.check(_lib.lib().rmq_server_submit(
            self._s, array_id, lr.ctypes.data, q, val.ctypes.data if want_val else None,
            idx.ctypes.data if want_idx else None, ctypes.byref(t)), "rmq_server_submit")
        self._live[t.value] = (lr, val, idx)
        return t.value, val, idx

    def _finish(self, fn) -> list:
        n = ctypes.c_uint32()
        _lib.check(fn(self._s, self._done_buf.ctypes.data, self._done_buf.size, ctypes.byref(n)),
                   fn.__name__)
        done = self._done_buf[: n.value].tolist()
        for t in done:
            self._live.pop(t, None)
        return done

    def poll(self) -> list:
        return self._finish(_lib.lib().rmq_server_poll)

    def drain(self) -> list:
        return self._finish(_lib.lib().rmq_server_drain)

    def stats(self) -> dict:
        st = _lib.ServerStats()
        _lib.check(_lib.lib().rmq_server_get_stats(self._s, ctypes.byref(st)), "rmq_server_get_stats")
        return {f: getattr(st, f) for f, _ in _lib.ServerStats._fields_}

    def close(self) -> None:
        if self._s.value is not None:
            _lib.lib().rmq_server_destroy(self._s)
            self._s = ctypes.c_void_p()
            self._live.clear()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
