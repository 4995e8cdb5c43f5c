"""ctypes loader for ``oracle/libknn_ref.so`` (the sequential C restatement of the
neighbour search in ``oracle/knn_ref.c``) -- TEST INFRASTRUCTURE ONLY."""
from __future__ import annotations

import ctypes
import os
import subprocess

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


def build() -> str:
    """Compile ``libknn_ref.so`` with ``make`` in this folder; returns its path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return os.path.join(_HERE, "libknn_ref.so")


def _lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "libknn_ref.so")
        if not os.path.exists(path):
            build()
        lib = ctypes.CDLL(path)
        for name in ("knn_ref_f32", "knn_ref_f64"):
            fn = getattr(lib, name)
            fn.restype = ctypes.c_int
            fn.argtypes = [ctypes.c_void_p] * 4 + [ctypes.c_int64] * 4 + [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]
        _LIB = lib
    return _LIB


def knn_ref(q, p, sigma, R, K: int, key: str = "aniso"):
    """Neighbour search by sequential selection; returns ``(dists, idxs)`` like
    ``oracle.plain.knn_sort``."""
    q, p, sigma, R = (t.detach().contiguous().cpu() for t in (q, p, sigma, R))
    B, Q, _ = q.shape
    P = p.shape[1]
    dt = q.dtype
    if dt not in (torch.float32, torch.float64):
        raise TypeError("float32 or float64 only")
    fn = _lib().knn_ref_f32 if dt == torch.float32 else _lib().knn_ref_f64
    dists = torch.empty(B, Q, K, dtype=dt)
    idxs = torch.empty(B, Q, K, dtype=torch.int64)
    mode = {"aniso": 0, "euclidean": 1}[key]
    rc = fn(q.data_ptr(), p.data_ptr(), sigma.data_ptr(), R.data_ptr(), B, Q, P, K, mode,
            dists.data_ptr(), idxs.data_ptr())
    if rc != 0:
        raise ValueError("K must satisfy 1 <= K <= P")
    return dists, idxs
