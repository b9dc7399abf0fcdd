"""ctypes binding of librmq.so (include/rmq.h, include/rmq_gen.h).

The shared library is built in-tree by ``make -C gpu-rmq_amd`` (``__graft_entry__.build()``).
Loading fails loudly when it is missing: there is no fallback path.
"""

from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.environ.get("RMQ_LIB") or os.path.join(PKG_DIR, "librmq.so")

RMQ_OK = 0
RMQ_EINVAL = -1
RMQ_ETHRESHOLD = -2
RMQ_ENOMEM = -3
RMQ_EDEVICE = -4
RMQ_EUNSUPPORTED = -5
RMQ_EAGAIN = -6

RMQ_ALG_GPU_BASE = 2
RMQ_ALG_XXX = 16
RMQ_ALG_VECTOR_LOAD = 20
RMQ_ALG_XXX_MULTI_LOAD = 24
RMQ_MAX_LEVELS = 20
RMQ_BUILD_DEFAULT_CG = 0x1
RMQ_BUILD_FULL = 0x2
RMQ_BUILD_NO_TOP_TABLE = 0x4
RMQ_BUILD_PACK_DESCENTS = 0x8
RMQ_BUILD_CHECK_QUERIES = 0x10
RMQ_BUILD_PAIRED = 0x20
RMQ_BUILD_BLOCK_SCANS = 0x40
RMQ_BUILD_NO_BLOCK_SCANS = 0x80


def RMQ_BUILD_CHUNK_LOG(k: int) -> int:
    """Chunk size (log2) of the 64-bit form for one build."""
    return (k & 0xFF) << 8
RMQ_VARIANT_LANE = 1
RMQ_VARIANT_LANE8 = 2
RMQ_VARIANT_LANE4 = 3
RMQ_VARIANT_GROUP4 = 4
RMQ_VARIANT_GROUP8 = 8
RMQ_VARIANT_GROUP16 = 16
RMQ_VARIANT_GROUP32 = 32
RMQ_VARIANT_WAVE = 64


RMQ_ADMIT_DEVICE = 0
RMQ_ADMIT_HOST = 1
RMQ_ADMIT_UNIFORM = 2


class ServerOptions(ctypes.Structure):
    """rmq_server_options (include/rmq_serve.h)."""

    _fields_ = [
        ("slots", ctypes.c_uint32),
        ("max_batch_queries", ctypes.c_uint32),
        ("reserve_bytes", ctypes.c_uint64),
        ("max_arrays", ctypes.c_uint32),
        ("max_batch_requests", ctypes.c_uint32),
        ("ring_queries", ctypes.c_uint64),
    ]


class ServerStats(ctypes.Structure):
    """rmq_server_stats (include/rmq_serve.h)."""

    _fields_ = [
        ("submitted", ctypes.c_uint64),
        ("completed", ctypes.c_uint64),
        ("batches", ctypes.c_uint64),
        ("queries", ctypes.c_uint64),
        ("resident_bytes", ctypes.c_uint64),
        ("resident", ctypes.c_uint32),
        ("in_flight", ctypes.c_uint32),
    ]


class BatchArray(ctypes.Structure):
    """rmq_batch_array (include/rmq_batch.h)."""

    _fields_ = [
        ("offset", ctypes.c_uint64),
        ("n", ctypes.c_uint32),
        ("t", ctypes.c_uint32),
        ("variant", ctypes.c_uint32),
        ("reserved", ctypes.c_uint32),
    ]

# Every symbol the public headers declare, with its ctypes signature.
_u32 = ctypes.c_uint32
_u64 = ctypes.c_uint64
_i32 = ctypes.c_int32
_vp = ctypes.c_void_p
_pu32 = ctypes.POINTER(ctypes.c_uint32)
_pu64 = ctypes.POINTER(ctypes.c_uint64)


class BuildOptions(ctypes.Structure):
    """rmq_build_options (include/rmq.h)."""

    _fields_ = [
        ("cg_size_log", ctypes.c_uint32),
        ("cg_amount_log", ctypes.c_uint32),
        ("group_lanes", ctypes.c_uint32),
        ("flags", ctypes.c_uint32),
    ]


SIGNATURES = {
    "rmq_build": (ctypes.c_int, [_vp, _u64, _u64, ctypes.c_int, _vp, ctypes.POINTER(_vp)]),
    "rmq_build_ex": (
        ctypes.c_int,
        [_vp, _u64, _u64, ctypes.c_int, ctypes.POINTER(BuildOptions), _vp, ctypes.POINTER(_vp)],
    ),
    "rmq_query": (ctypes.c_int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "rmq_free": (None, [_vp]),
    "rmq_get_metadata": (ctypes.c_int, [_vp, _pu32, _pu64, _pu64, _pu64, _pu32]),
    "rmq_plan": (
        ctypes.c_int,
        [_u64, _u64, ctypes.c_int, ctypes.POINTER(BuildOptions), _pu32, _pu64, _pu64, _pu64, _pu32],
    ),
    "rmq_read_levels": (ctypes.c_int, [_vp, _vp, _vp, _vp]),
    "rmq_gpu_base": (ctypes.c_int, [_vp, _u64, _vp, _u64, _vp, _vp, _vp]),
    "rmq_build_each": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.POINTER(BuildOptions),
                                      _u32, _vp, _vp]),
    "rmq_query_each": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _u32, _vp]),
    "rmq_status_str": (ctypes.c_char_p, [ctypes.c_int]),
    "rmq_version": (ctypes.c_char_p, []),
    "rmq_batch_build": (ctypes.c_int, [_vp, _vp, _u32, _vp, ctypes.POINTER(_vp)]),
    "rmq_batch_plan_queries": (ctypes.c_int, [_vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "rmq_batch_query": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp]),
    "rmq_batch_query_noalloc": (ctypes.c_int, [_vp, _vp, _u64, _vp, _vp, _vp, _vp]),
    "rmq_batch_capture": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.POINTER(_vp)]),
    "rmq_graph_launch": (ctypes.c_int, [_vp, _vp]),
    "rmq_graph_free": (None, [_vp]),
    "rmq_batch_plan_free": (None, [_vp]),
    "rmq_batch_free": (None, [_vp]),
    "rmq_batch_level_bytes": (_u64, [_vp]),
    "rmq_batch_read_levels": (ctypes.c_int, [_vp, _u32, _vp, _vp, _vp]),
    "rmq_fill_uniform": (ctypes.c_int, [_vp, _u64, _u64, _u64, _vp]),
    "rmq_server_create": (ctypes.c_int, [ctypes.POINTER(ServerOptions), ctypes.POINTER(_vp)]),
    "rmq_server_destroy": (None, [_vp]),
    "rmq_server_admit": (ctypes.c_int, [_vp, ctypes.c_int, _vp, _u64, _u64, _u32, _u32,
                                        ctypes.POINTER(ctypes.c_int32)]),
    "rmq_server_evict": (ctypes.c_int, [_vp, ctypes.c_int32]),
    "rmq_server_submit": (ctypes.c_int, [_vp, ctypes.c_int32, _vp, _u64, _vp, _vp,
                                         ctypes.POINTER(ctypes.c_uint64)]),
    "rmq_server_poll": (ctypes.c_int, [_vp, _vp, _u32, ctypes.POINTER(ctypes.c_uint32)]),
    "rmq_server_drain": (ctypes.c_int, [_vp, _vp, _u32, ctypes.POINTER(ctypes.c_uint32)]),
    "rmq_server_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(ServerStats)]),
    "rmq_gen_array": (None, [_u64, _i32, _vp]),
    "rmq_gen_queries": (ctypes.c_int, [_u64, _i32, _u64, _i32, ctypes.c_int, _vp]),
    "rmq_gen_queries64": (ctypes.c_int, [_u64, _i32, _u64, _i32, ctypes.c_int, _vp]),
    "rmq_query64": (ctypes.c_int, [_vp, _vp, _u64, _vp, _vp, _vp]),
    "rmq_gpu_base64": (ctypes.c_int, [_vp, _u64, _vp, _u64, _vp, _vp, _vp]),
    "rmq_gen_sample_indices": (_u64, [_u64, _u64, _i32, _vp]),
}

_lib = None


class RMQError(RuntimeError):
    """A non-zero rmq_status from the library."""

    def __init__(self, status: int, where: str):
        self.status = status
        super().__init__(f"{where}: {status_str(status)} (status {status})")


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"librmq.so not found at {LIB_PATH}; build it with `make -C gpu-rmq_amd` "
                "(or __graft_entry__.build())"
            )
        # One HIP runtime per process: torch bundles its own libamdhip64 (SONAME libamdhip64.so.7)
        # and loads it under the file name libamdhip64.so through its RPATH, while librmq.so asks
        # for libamdhip64.so.7. Loaded after torch, librmq.so binds to torch's copy by SONAME;
        # loaded first, it would bring in the system copy and torch would then load its own as a
        # second runtime, with neither seeing the other's streams or allocations. So torch (the
        # package's source of device memory and streams) is imported first when it is installed.
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        handle = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(handle, name)
            fn.restype = res
            fn.argtypes = args
        _lib = handle
    return _lib


def status_str(status: int) -> str:
    return lib().rmq_status_str(status).decode()


def check(status: int, where: str) -> None:
    if status != RMQ_OK:
        raise RMQError(status, where)


def version() -> str:
    return lib().rmq_version().decode()
