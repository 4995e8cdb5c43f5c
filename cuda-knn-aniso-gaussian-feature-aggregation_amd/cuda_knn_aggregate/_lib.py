"""ctypes binding of ``libaknn.so`` (C ABI declared in ``include/aknn.h``).

The library is loaded lazily on first use.  It links ``libamdhip64.so.7`` by soname;
torch is imported first so the process already holds torch's HIP runtime and the
library binds to that same runtime (one HIP context, shared device pointers and
streams).  There is no fallback: a missing or unloadable library raises.
"""
from __future__ import annotations

import ctypes
import os
import threading

import torch  # noqa: F401  (loads the HIP runtime the library binds to)

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libaknn.so")

ABI_VERSION = 2

AKNN_OK = 0
AKNN_E_INVALID = -1
AKNN_E_WORKSPACE = -2
AKNN_E_LAUNCH = -3
AKNN_E_UNSUPPORTED = -4

AKNN_F32 = 0
AKNN_F64 = 1
KEYS = {"aniso": 0, "euclidean": 1}

NEED_DQ = 1
NEED_DP = 2
NEED_DF = 4
NEED_DSIGMA = 8
NEED_DR = 16
NEED_DGRAD_F_OUT = 32
SECOND_FULL = 64
GRAD_BROADCAST = 128  # grad_f_out given as one row per cloud ([B, F])
NEED_DGRAD_W_OUT = 256  # aknn_backward2_w: d grad_w_out
CSR_QUERY_ORDER = 512  # the CSR numbers queries by their position in the cache's query order
NEED_POINTS = NEED_DP | NEED_DF | NEED_DSIGMA | NEED_DR

EXPORTS = (
    "aknn_abi_version",
    "aknn_last_error",
    "aknn_trace_enable",
    "aknn_trace_collect",
    "aknn_forward_workspace_size",
    "aknn_forward",
    "aknn_knn",
    "aknn_aggregate",
    "aknn_csr_workspace_size",
    "aknn_build_csr",
    "aknn_build_csr_ordered",
    "aknn_backward_workspace_size",
    "aknn_backward",
    "aknn_backward2",
    "aknn_backward2_full",
    "aknn_backward_w",
    "aknn_backward2_w",
    "aknn_index_size",
    "aknn_index_workspace_size",
    "aknn_index_build",
    "aknn_lookup_workspace_size",
    "aknn_lookup_forward",
    "aknn_iso_workspace_size",
    "aknn_iso_forward",
    "aknn_point_dots",
    "aknn_forward_noalloc",
    "aknn_copy",
    "aknn_host_device_pointer",
)


class Dims(ctypes.Structure):
    """``aknn_dims`` (include/aknn.h)."""

    _fields_ = [
        ("B", ctypes.c_int64),
        ("Q", ctypes.c_int64),
        ("P", ctypes.c_int64),
        ("F", ctypes.c_int64),
        ("K", ctypes.c_int64),
        ("dtype", ctypes.c_int32),
        ("key", ctypes.c_int32),
        ("q_off", ctypes.c_void_p),
        ("p_off", ctypes.c_void_p),
        ("q_off_host", ctypes.c_void_p),
        ("p_off_host", ctypes.c_void_p),
    ]


class Cache(ctypes.Structure):
    """``aknn_cache`` (include/aknn.h)."""

    _fields_ = [("query_order", ctypes.c_void_p), ("point_order", ctypes.c_void_p),
                ("point_records", ctypes.c_void_p), ("point_dots", ctypes.c_void_p)]


class CopyItem(ctypes.Structure):
    """``aknn_copy_item`` (include/aknn.h)."""

    _fields_ = [("dst", ctypes.c_void_p), ("src", ctypes.c_void_p), ("bytes", ctypes.c_uint64)]


COPY_MAX_ITEMS = 8


class Tune(ctypes.Structure):
    """``aknn_tune`` (include/aknn.h)."""

    _fields_ = [("queries_per_lane", ctypes.c_int32), ("point_split", ctypes.c_int32),
                ("search", ctypes.c_int32), ("sub_records", ctypes.c_int32), ("stats", ctypes.c_void_p),
                ("index", ctypes.c_void_p), ("super_tiles", ctypes.c_int32), ("waves_per_eu", ctypes.c_int32),
                ("mu_classes", ctypes.c_int32), ("persistent", ctypes.c_int32)]


SEARCH = {"auto": 0, "brute": 1, "tiles": 2, "wave": 3}


_lock = threading.Lock()
_lib = None

_P = ctypes.c_void_p
_DIMS = ctypes.POINTER(Dims)
_TUNE = ctypes.POINTER(Tune)
_CACHE = ctypes.POINTER(Cache)
_SIGS = {
    "aknn_abi_version": (ctypes.c_int, []),
    "aknn_last_error": (ctypes.c_char_p, []),
    "aknn_trace_enable": (None, [ctypes.c_int]),
    "aknn_trace_collect": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t, ctypes.POINTER(ctypes.c_float), ctypes.c_int]),
    "aknn_forward_workspace_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_forward": (ctypes.c_int, [_DIMS, _TUNE] + [_P] * 9 + [_P, ctypes.c_size_t, _CACHE, _P]),
    "aknn_knn": (ctypes.c_int, [_DIMS, _TUNE] + [_P] * 6 + [_P, ctypes.c_size_t, _P]),
    "aknn_aggregate": (ctypes.c_int, [_DIMS] + [_P] * 6 + [_P] * 3 + [_P]),
    "aknn_csr_workspace_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_build_csr": (ctypes.c_int, [_DIMS, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    "aknn_build_csr_ordered": (ctypes.c_int, [_DIMS, _P, _P, _P, _P, _P, ctypes.c_size_t, _P]),
    "aknn_backward_workspace_size": (ctypes.c_size_t, [_DIMS, ctypes.c_uint32]),
    "aknn_backward": (ctypes.c_int, [_DIMS, ctypes.c_uint32] + [_P] * 9 + [_CACHE, _P, ctypes.c_size_t]
                      + [_P] * 5 + [_P]),
    "aknn_backward2": (ctypes.c_int, [_DIMS, ctypes.c_uint32] + [_P] * 10 + [_CACHE, _P, ctypes.c_size_t]
                       + [_P] * 6 + [_P]),
    "aknn_backward2_full": (ctypes.c_int, [_DIMS, ctypes.c_uint32] + [_P] * 14 + [_CACHE, _P, ctypes.c_size_t]
                            + [_P] * 6 + [_P]),
    "aknn_backward_w": (ctypes.c_int, [_DIMS, ctypes.c_uint32] + [_P] * 10 + [_CACHE, _P, ctypes.c_size_t]
                        + [_P] * 5 + [_P]),
    "aknn_backward2_w": (ctypes.c_int, [_DIMS, ctypes.c_uint32] + [_P] * 15 + [_CACHE, _P, ctypes.c_size_t]
                         + [_P] * 7 + [_P]),
    "aknn_index_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_index_workspace_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_index_build": (ctypes.c_int, [_DIMS, _P, _P, _P, _P, ctypes.c_size_t, _P, ctypes.c_size_t, _P]),
    "aknn_lookup_workspace_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_lookup_forward": (ctypes.c_int, [_DIMS] + [_P] * 6 + [ctypes.c_int64] + [_P] * 4
                            + [_P, ctypes.c_size_t, _P]),
    "aknn_iso_workspace_size": (ctypes.c_size_t, [_DIMS]),
    "aknn_iso_forward": (ctypes.c_int, [_DIMS, _TUNE] + [_P] * 4 + [_P] * 4 + [_P, ctypes.c_size_t, _P]),
    "aknn_point_dots": (ctypes.c_int, [_DIMS] + [_P] * 6 + [_P]),
    "aknn_forward_noalloc": (ctypes.c_int, [_DIMS, _TUNE] + [_P] * 9 + [_CACHE, _P]),
    "aknn_copy": (ctypes.c_int, [ctypes.POINTER(CopyItem), ctypes.c_int, _P]),
    "aknn_host_device_pointer": (ctypes.c_int, [_P, ctypes.POINTER(ctypes.c_void_p)]),
}
NOALLOC_SCRATCH_BYTES = 64 << 20


def load(path: str | None = None):
    """Load (once) and return the ctypes handle of ``libaknn.so``."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        path = path or LIB_PATH
        if not os.path.exists(path):
            raise RuntimeError(
                f"libaknn.so not found at {path}: build it first (__graft_entry__.build() or "
                "python cuda_knn_aggregate/_build.py)")
        lib = ctypes.CDLL(path)
        for name, (res, args) in _SIGS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        ver = lib.aknn_abi_version()
        if ver != ABI_VERSION:
            raise RuntimeError(f"libaknn.so ABI version {ver} != expected {ABI_VERSION}")
        _lib = lib
        return lib


def check(rc: int) -> None:
    """Raise ``RuntimeError`` with the library's message for a non-zero return code."""
    if rc != AKNN_OK:
        msg = load().aknn_last_error()
        raise RuntimeError((msg.decode() if msg else "libaknn error") + f" (code {rc})")


def host_device_pointer(t) -> int:
    """Device address of a pinned host tensor (``pin_memory()``), for :func:`copy`."""
    out = ctypes.c_void_p()
    check(load().aknn_host_device_pointer(ctypes.c_void_p(t.data_ptr()), ctypes.byref(out)))
    return out.value


def copy(items, stream) -> None:
    """Kernel-based copies ``[(dst_address, src_address, bytes), ...]`` on ``stream`` (a
    hipStream_t as int); either side may be device memory or mapped pinned host memory."""
    lib = load()
    for s0 in range(0, len(items), COPY_MAX_ITEMS):
        part = items[s0:s0 + COPY_MAX_ITEMS]
        arr = (CopyItem * len(part))(*[CopyItem(d, s, n) for d, s, n in part])
        check(lib.aknn_copy(arr, len(part), stream))


def ptr(t) -> int | None:
    """Device address of a tensor (or None)."""
    return None if t is None else t.data_ptr()


class trace:
    """Context manager recording the library's kernel launches with hipEvents.

    ``with trace() as t: ...`` then ``t.records`` is a list of ``(name, ms)`` in launch
    order and ``t.summary()`` maps name -> (launches, total ms).
    """

    def __init__(self):
        self.records = []

    def __enter__(self):
        lib = load()
        lib.aknn_trace_collect(None, 0, None, 0)  # drop stale records
        lib.aknn_trace_enable(1)
        return self

    def __exit__(self, *exc):
        lib = load()
        lib.aknn_trace_enable(0)
        self.records = collect()
        return False

    def summary(self):
        out = {}
        for name, ms in self.records:
            n, tot = out.get(name, (0, 0.0))
            out[name] = (n + 1, tot + ms)
        return out


def collect(max_records: int = 1 << 16):
    """Wait for and return the traced launches as ``[(name, ms), ...]``."""
    lib = load()
    names = ctypes.create_string_buffer(max_records * 24)
    ms = (ctypes.c_float * max_records)()
    n = lib.aknn_trace_collect(names, len(names), ms, max_records)
    raw = names.raw.split(b"\0")
    return [(raw[i].decode(), float(ms[i])) for i in range(n)]
