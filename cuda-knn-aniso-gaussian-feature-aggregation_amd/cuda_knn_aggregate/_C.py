"""Tensor-level entry points, one per export of the reference's pybind11 module ``_C``.

Reference: csrc/knn_aggregate.cpp:178-185 exports ``knn_aggregate_aniso_forward``,
``knn_aggregate_aniso_backward`` and ``knn_aggregate_aniso_backward_2nd`` (plus the
isotropic, lookup and precomputed variants, which are not on this path).  The
functions below keep those names, argument orders and return lists, check their
inputs like ``CHECK_INPUT`` (csrc/knn_aggregate.h:6-8: device and contiguity, raising
``RuntimeError``) plus dtype, shape and ``K`` range, allocate the outputs with torch
and run the HIP library on the current torch stream.  CPU tensors are rejected:
there is no CPU path.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import torch

from . import _lib

__all__ = [
    "knn_aggregate_aniso_forward",
    "knn_precomputed_aggregate_aniso_forward",
    "knn_lookup_aggregate_aniso_forward",
    "knn_aggregate_forward",
    "knn_aggregate_aniso_backward",
    "knn_aggregate_aniso_backward_2nd",
    "knn_aniso",
    "build_csr",
    "build_index",
    "make_dims",
    "new_cache",
    "broadcast_rows",
    "upstream",
    "point_dots",
]

_DTYPES = {torch.float32: _lib.AKNN_F32, torch.float64: _lib.AKNN_F64}


# ----------------------------------------------------------------------------- checks
def _check_input(t: torch.Tensor, name: str) -> None:
    """``CHECK_INPUT`` (csrc/knn_aggregate.h:6-8): a contiguous GPU tensor."""
    if not isinstance(t, torch.Tensor):
        raise RuntimeError(f"{name} must be a tensor")
    if not t.is_cuda:
        raise RuntimeError(f"{name} must be a CUDA tensor")
    if not t.is_contiguous():
        raise RuntimeError(f"{name} must be contiguous")


def _check_float(ts: Sequence[torch.Tensor], names: Sequence[str]) -> torch.dtype:
    dt = ts[0].dtype
    if dt not in _DTYPES:
        raise RuntimeError(f"{names[0]} must be float32 or float64 (got {dt})")
    for t, n in zip(ts, names):
        if t.dtype != dt:
            raise RuntimeError(f"{n} must have dtype {dt} (got {t.dtype})")
    dev = ts[0].device
    for t, n in zip(ts, names):
        if t.device != dev:
            raise RuntimeError(f"{n} must be on {dev} (got {t.device})")
    return dt


def _check_shapes(q, p, f, sigma, R):
    if q.dim() != 3 or q.shape[2] != 3:
        raise RuntimeError(f"q must have shape [B, Q, 3] (got {tuple(q.shape)})")
    B = q.shape[0]
    if p.dim() != 3 or p.shape[0] != B or p.shape[2] != 3:
        raise RuntimeError(f"p must have shape [B, P, 3] (got {tuple(p.shape)})")
    P = p.shape[1]
    if f.dim() != 3 or f.shape[0] != B or f.shape[1] != P:
        raise RuntimeError(f"f must have shape [B, P, F] (got {tuple(f.shape)})")
    if sigma.dim() != 3 or tuple(sigma.shape) != (B, P, 3):
        raise RuntimeError(f"sigma must have shape [B, P, 3] (got {tuple(sigma.shape)})")
    if R.dim() != 4 or tuple(R.shape) != (B, P, 3, 3):
        raise RuntimeError(f"R must have shape [B, P, 3, 3] (got {tuple(R.shape)})")
    if f.shape[2] < 1:
        raise RuntimeError("f must have at least one feature")
    return B, q.shape[1], P, f.shape[2]


def _check_k(K, P) -> int:
    if isinstance(K, bool) or not isinstance(K, int):
        try:
            K = int(K)
        except Exception as exc:  # noqa: BLE001
            raise RuntimeError(f"K must be an int (got {K!r})") from exc
    if K < 1:
        raise RuntimeError(f"K must be >= 1 (got {K})")
    if K > P:
        raise RuntimeError(f"K ({K}) must not exceed the number of points P ({P})")
    if K > 64:
        raise RuntimeError(f"K must be <= 64 (got {K})")
    return K


def make_dims(B, Q, P, F, K, dtype, key="aniso", offsets=None) -> _lib.Dims:
    """Fill an ``aknn_dims``.  ``offsets`` = (q_off_dev, p_off_dev, q_off_host, p_off_host)
    for ragged batches (int64 tensors of B+1 entries; the host copies must stay alive)."""
    d = _lib.Dims()
    d.B, d.Q, d.P, d.F, d.K = int(B), int(Q), int(P), int(F), int(K)
    d.dtype = _DTYPES[dtype]
    if key not in _lib.KEYS:
        raise RuntimeError(f"key must be one of {sorted(_lib.KEYS)} (got {key!r})")
    d.key = _lib.KEYS[key]
    if offsets is not None:
        qd, pd, qh, ph = offsets
        d.q_off, d.p_off = qd.data_ptr(), pd.data_ptr()
        d.q_off_host, d.p_off_host = qh.data_ptr(), ph.data_ptr()
    return d


def _stream(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _tune(tune) -> Optional[_lib.Tune]:
    if tune is None:
        return None
    t = _lib.Tune()
    t.queries_per_lane = int(tune.get("queries_per_lane", 0))
    t.point_split = int(tune.get("point_split", 0))
    t.search = _lib.SEARCH[tune.get("search", "auto")]
    t.sub_records = int(tune.get("sub_records", 0))
    t.super_tiles = int(tune.get("super_tiles", 0))
    t.waves_per_eu = int(tune.get("waves_per_eu", 0))
    t.mu_classes = int(tune.get("mu_classes", 0))
    t.persistent = int(tune.get("persistent", 0))
    st = tune.get("stats")  # optional int64 device tensor of 5 counters
    t.stats = st.data_ptr() if st is not None else None
    ix = tune.get("index")  # optional point index from build_index (same points)
    t.index = ix.data_ptr() if ix is not None else None
    return t


def _tp(t: Optional[_lib.Tune]):
    return ctypes.byref(t) if t is not None else None


def new_cache(B, Q, P, dtype, device):
    """Tensors of an ``aknn_cache`` for a [B,Q] x [B,P] batch: spatial orders of the
    query and point rows and packed point records (filled by the forward)."""
    return (torch.empty(B * Q, dtype=torch.int32, device=device),
            torch.empty(B * P, dtype=torch.int32, device=device),
            torch.empty(B * P * 16, dtype=dtype, device=device))


def _cache_struct(cache):
    """``aknn_cache`` from (query_order, point_order, point_records[, point_dots])."""
    if cache is None:
        return None
    s = _lib.Cache()
    ptrs = [t.data_ptr() if t is not None else None for t in cache] + [None] * (4 - len(cache))
    s.query_order, s.point_order, s.point_records, s.point_dots = ptrs
    return s


# ----------------------------------------------------------------------------- upstream
def broadcast_rows(g, B: int, Q: int, F: int):
    """``[B, F]`` rows of an upstream gradient ``g`` [B, Q, F] that repeats one row along the
    query axis as a stride-0 view (``v.expand_as(f_out)``, as in the Eikonal read-out), or
    ``None`` for any other layout.  The reference materialises such an upstream with
    ``.contiguous()`` (cuda_knn_aggregate/__init__.py:107); the library takes its rows
    instead (``AKNN_GRAD_BROADCAST``) with the same results."""
    if not isinstance(g, torch.Tensor) or g.dim() != 3 or tuple(g.shape) != (B, Q, F) or Q < 1:
        return None
    if g.stride(1) != 0 or (g.stride(2) != 1 and F > 1):
        return None
    return g[:, 0, :].contiguous()


def upstream(g):
    """An upstream gradient as the backward entry points take it: broadcast views stay
    views (their rows are read as such), anything else is made contiguous."""
    if g is None:
        return None
    if g.dim() == 3 and broadcast_rows(g, *g.shape) is not None:
        return g
    return g.contiguous()


def _gcols(g, rows, s0, s1):
    """Columns [s0, s1) of an upstream gradient, keeping a broadcast upstream broadcast."""
    if rows is None:
        return _cols(g, s0, s1)
    B, Q = g.shape[0], g.shape[1]
    return rows[:, None, s0:s1].expand(B, Q, s1 - s0)


def _cp(s):
    return ctypes.byref(s) if s is not None else None


# ----------------------------------------------------------------------------- forward
def _feature_slabs(F: int, dt: torch.dtype):
    """Column ranges ``[s0, s1)`` splitting a feature row wider than the library takes in
    one call, or ``None`` when ``F`` fits.  The library holds a row in registers: up to
    128 times the widest vector (at most 16 B) that divides the width, so 512 (FP32,
    width divisible by 4), 256 (even) or 128.  Every term of the forward and of both
    backward orders is either local to a feature column or linear in the row dot
    products (``sum_F g f`` and the like), so per-range calls add up to the whole row:
    column outputs are placed side by side, geometric gradients summed in range order.
    """
    def cap(n: int) -> int:
        v = 4 if (dt == torch.float32 and n % 4 == 0) else (2 if n % 2 == 0 else 1)
        return 128 * v

    if F <= cap(F):
        return None
    width = 512 if dt == torch.float32 else 256
    out, s0 = [], 0
    while F - s0 > width:
        out.append((s0, s0 + width))
        s0 += width
    if F - s0 <= cap(F - s0):
        out.append((s0, F))
    else:
        out += [(a, min(a + 128, F)) for a in range(s0, F, 128)]
    return out


def _cols(t, s0, s1):
    return None if t is None else t[..., s0:s1].contiguous()


def knn_aggregate_aniso_forward(q, p, f, sigma, R, K, *, key: str = "aniso", tune=None, cache=None):
    """Neighbour search + aggregation -> ``[f_out, w_out, dists, idxs]``.

    Reference: ``knn_aggregate_aniso_forward`` (csrc/knn_aggregate.cpp:26-47,
    csrc/forward.cu:286-338).  ``f_out`` [B,Q,F], ``w_out``/``dists`` [B,Q,K],
    ``idxs`` [B,Q,K] int64; neighbours ordered by (key, point index).  ``cache`` (from
    :func:`new_cache`) is filled for the backward passes; results never depend on it.
    """
    for t, n in zip((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R")):
        _check_input(t, n)
    dt = _check_float((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R"))
    B, Q, P, F = _check_shapes(q, p, f, sigma, R)
    K = _check_k(K, P)
    slabs = _feature_slabs(F, dt)
    if slabs is not None:  # one search, then the other column ranges on its neighbours
        (s0, s1), rest = slabs[0], slabs[1:]
        f_out = torch.empty(B, Q, F, dtype=dt, device=q.device)
        first, w_out, dists, idxs = knn_aggregate_aniso_forward(q, p, _cols(f, s0, s1), sigma, R, K, key=key,
                                                                tune=tune, cache=cache)
        f_out[..., s0:s1] = first
        for s0, s1 in rest:
            f_out[..., s0:s1] = knn_precomputed_aggregate_aniso_forward(q, p, _cols(f, s0, s1), sigma, R, idxs, K,
                                                                        key=key, check=False)[0]
        return [f_out, w_out, dists, idxs]
    dev = q.device
    f_out = torch.empty(B, Q, F, dtype=dt, device=dev)
    w_out = torch.empty(B, Q, K, dtype=dt, device=dev)
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    idxs = torch.empty(B, Q, K, dtype=torch.int64, device=dev)
    if B * Q == 0:
        return [f_out, w_out, dists, idxs]
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt, key)
    ws_bytes = lib.aknn_forward_workspace_size(ctypes.byref(dims))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    tn = _tune(tune)
    cs = _cache_struct(cache)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_forward(ctypes.byref(dims), _tp(tn), q.data_ptr(), p.data_ptr(),
                                    f.data_ptr(), sigma.data_ptr(), R.data_ptr(),
                                    f_out.data_ptr(), w_out.data_ptr(), dists.data_ptr(),
                                    idxs.data_ptr(), ws.data_ptr(), ws_bytes, _cp(cs), _stream(dev)))
    return [f_out, w_out, dists, idxs]


def knn_precomputed_aggregate_aniso_forward(q, p, f, sigma, R, knn_idxs, K, *, key: str = "aniso",
                                            check: bool = True):
    """Aggregation given neighbour indices -> ``[f_out, w_out, dists]``.

    Reference: ``knn_precomputed_aggregate_aniso_forward`` (csrc/knn_aggregate.cpp:76-99,
    csrc/forward.cu:186-230,398-450).  ``knn_idxs`` [B,Q,K] int64 must index [0, P);
    it is range-checked here (one device reduction and a host synchronisation) because
    the kernel trusts it; ``check=False`` skips that for indices known to be in range.
    """
    for t, n in zip((q, p, f, sigma, R, knn_idxs), ("q", "p", "f", "sigma", "R", "knn_idxs")):
        _check_input(t, n)
    dt = _check_float((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R"))
    B, Q, P, F = _check_shapes(q, p, f, sigma, R)
    if knn_idxs.dtype != torch.int64 or knn_idxs.dim() != 3 or tuple(knn_idxs.shape[:2]) != (B, Q):
        raise RuntimeError(f"knn_idxs must be int64 [B, Q, K] (got {tuple(knn_idxs.shape)}, {knn_idxs.dtype})")
    if int(knn_idxs.shape[2]) != int(K):
        raise RuntimeError(f"knn_idxs has {knn_idxs.shape[2]} neighbours but K = {K}")
    K = _check_k(K, P)
    slabs = _feature_slabs(F, dt)
    if slabs is not None:
        f_out = torch.empty(B, Q, F, dtype=dt, device=q.device)
        if check:
            _check_range(knn_idxs, P, "knn_idxs")
        for i, (s0, s1) in enumerate(slabs):
            o = knn_precomputed_aggregate_aniso_forward(q, p, _cols(f, s0, s1), sigma, R, knn_idxs, K, key=key,
                                                        check=False)
            f_out[..., s0:s1] = o[0]
            if i == 0:
                w_out, dists = o[1], o[2]
        return [f_out, w_out, dists]
    dev = q.device
    f_out = torch.empty(B, Q, F, dtype=dt, device=dev)
    w_out = torch.empty(B, Q, K, dtype=dt, device=dev)
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    if B * Q == 0:
        return [f_out, w_out, dists]
    if check:
        _check_range(knn_idxs, P, "knn_idxs")
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt, key)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_aggregate(ctypes.byref(dims), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                                      sigma.data_ptr(), R.data_ptr(), knn_idxs.data_ptr(),
                                      f_out.data_ptr(), w_out.data_ptr(), dists.data_ptr(),
                                      _stream(dev)))
    return [f_out, w_out, dists]


def knn_lookup_aggregate_aniso_forward(q, p, f, sigma, R, knn_table, K, *, check: bool = True):
    """Table-lookup forward -> ``[f_out, w_out, dists, idxs]``.

    Reference: ``knn_lookup_aggregate_aniso_forward`` (csrc/knn_aggregate.cpp:50-73,
    csrc/forward.cu:113-183,341-395).  Each query's Euclidean nearest point ``n1`` (lowest
    index at the minimum; 0 unless that minimum is below 1e10) picks row ``n1`` of
    ``knn_table`` [P, >=K] (int64, shared by the batches); its first K entries are the
    neighbours, in table order, weighted with the anisotropic Gaussians.  The table is
    range-checked here (one device reduction and a host synchronisation) because the
    kernels trust it; ``check=False`` skips that for a table known to be in range.
    """
    for t, n in zip((q, p, f, sigma, R, knn_table), ("q", "p", "f", "sigma", "R", "knn_table")):
        _check_input(t, n)
    dt = _check_float((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R"))
    B, Q, P, F = _check_shapes(q, p, f, sigma, R)
    if knn_table.dtype != torch.int64 or knn_table.dim() != 2 or int(knn_table.shape[0]) != P:
        raise RuntimeError(f"knn_table must be int64 [P, >=K] (got {tuple(knn_table.shape)}, {knn_table.dtype})")
    if knn_table.device != q.device:
        raise RuntimeError("knn_table must be on the device of q")
    K = _check_k(K, P)
    if int(knn_table.shape[1]) < K:
        raise RuntimeError(f"knn_table has {knn_table.shape[1]} columns but K = {K}")
    slabs = _feature_slabs(F, dt)
    if slabs is not None:  # one lookup, then the other column ranges on its neighbours
        (s0, s1), rest = slabs[0], slabs[1:]
        f_out = torch.empty(B, Q, F, dtype=dt, device=q.device)
        first, w_out, dists, idxs = knn_lookup_aggregate_aniso_forward(q, p, _cols(f, s0, s1), sigma, R, knn_table, K,
                                                                       check=check)
        f_out[..., s0:s1] = first
        for s0, s1 in rest:
            f_out[..., s0:s1] = knn_precomputed_aggregate_aniso_forward(q, p, _cols(f, s0, s1), sigma, R, idxs, K,
                                                                        check=False)[0]
        return [f_out, w_out, dists, idxs]
    dev = q.device
    f_out = torch.empty(B, Q, F, dtype=dt, device=dev)
    w_out = torch.empty(B, Q, K, dtype=dt, device=dev)
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    idxs = torch.empty(B, Q, K, dtype=torch.int64, device=dev)
    if B * Q == 0:
        return [f_out, w_out, dists, idxs]
    if check:
        _check_range(knn_table[:, :K], P, "knn_table entries")
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt)
    ws_bytes = lib.aknn_lookup_workspace_size(ctypes.byref(dims))
    if ws_bytes == 0:
        raise RuntimeError(f"aknn_lookup_workspace_size failed: {lib.aknn_last_error().decode()}")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_lookup_forward(ctypes.byref(dims), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                                           sigma.data_ptr(), R.data_ptr(), knn_table.data_ptr(),
                                           int(knn_table.stride(0)), f_out.data_ptr(), w_out.data_ptr(),
                                           dists.data_ptr(), idxs.data_ptr(), ws.data_ptr(), ws_bytes,
                                           _stream(dev)))
    return [f_out, w_out, dists, idxs]


def knn_aggregate_forward(q, p, f, sigma, K, *, tune=None):
    """Isotropic forward -> ``[f_out, w_out [B,Q,1], dists, idxs]``.

    Reference: ``knn_aggregate_forward`` (csrc/knn_aggregate.cpp:5-23, csrc/forward.cu:5-53,
    launcher :233-283; exported but without a Python wrapper or backward there).  The K
    nearest points by Euclidean distance, ``dists`` = squared distances,
    ``w = exp(-dists * sigma[..., 0])``, ``f_out = sum w f``, ``w_out`` = the weight sum.
    ``sigma`` is [B, P, C] with C >= 1 (only channel 0 is read, as in the reference).
    """
    for t, n in zip((q, p, f, sigma), ("q", "p", "f", "sigma")):
        _check_input(t, n)
    dt = _check_float((q, p, f, sigma), ("q", "p", "f", "sigma"))
    if q.dim() != 3 or q.shape[2] != 3 or p.dim() != 3 or p.shape[2] != 3:
        raise RuntimeError("q and p must be [B, Q, 3] and [B, P, 3]")
    B, Q, P = q.shape[0], q.shape[1], p.shape[1]
    if p.shape[0] != B or f.dim() != 3 or tuple(f.shape[:2]) != (B, P) or f.shape[2] < 1:
        raise RuntimeError(f"f must be [B, P, F] with F >= 1 (got {tuple(f.shape)})")
    if sigma.dim() != 3 or tuple(sigma.shape[:2]) != (B, P) or sigma.shape[2] < 1:
        raise RuntimeError(f"sigma must be [B, P, C] with C >= 1 (got {tuple(sigma.shape)})")
    F = f.shape[2]
    K = _check_k(K, P)
    slabs = _feature_slabs(F, dt)
    if slabs is not None:  # each column range searches again (same neighbours, same weights)
        f_out = torch.empty(B, Q, F, dtype=dt, device=q.device)
        for i, (s0, s1) in enumerate(slabs):
            o = knn_aggregate_forward(q, p, _cols(f, s0, s1), sigma, K, tune=tune)
            f_out[..., s0:s1] = o[0]
            if i == 0:
                w_out, dists, idxs = o[1], o[2], o[3]
        return [f_out, w_out, dists, idxs]
    dev = q.device
    scale = sigma[..., 0].contiguous()
    f_out = torch.empty(B, Q, F, dtype=dt, device=dev)
    w_out = torch.empty(B, Q, 1, dtype=dt, device=dev)
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    idxs = torch.empty(B, Q, K, dtype=torch.int64, device=dev)
    if B * Q == 0:
        return [f_out, w_out, dists, idxs]
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt, "euclidean")
    ws_bytes = lib.aknn_iso_workspace_size(ctypes.byref(dims))
    if ws_bytes == 0:
        raise RuntimeError(f"aknn_iso_workspace_size failed: {lib.aknn_last_error().decode()}")
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    tn = _tune(tune)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_iso_forward(ctypes.byref(dims), _tp(tn), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                                        scale.data_ptr(), f_out.data_ptr(), w_out.data_ptr(), dists.data_ptr(),
                                        idxs.data_ptr(), ws.data_ptr(), ws_bytes, _stream(dev)))
    return [f_out, w_out, dists, idxs]


def build_index(p, sigma, R, *, key: str = "aniso", offsets=None):
    """Point index of a resident cloud batch (``p`` [B,P,3], ``sigma``, ``R``; or packed
    ragged rows with ``offsets`` = (q_off, p_off, q_off_host, p_off_host)) as a uint8
    device tensor: pass it as ``tune={"index": ...}`` to the forward / search of the same
    points and the tiled search skips its per-call point preparation (same results)."""
    for t, n in zip((p, sigma, R), ("p", "sigma", "R")):
        _check_input(t, n)
    dt = _check_float((p, sigma, R), ("p", "sigma", "R"))
    if offsets is None:
        if p.dim() != 3 or p.shape[2] != 3:
            raise RuntimeError(f"p must have shape [B, P, 3] (got {tuple(p.shape)})")
        B, P = p.shape[0], p.shape[1]
        dims = make_dims(B, 1, P, 1, 1, dt, key)
    else:
        B = int(offsets[2].numel()) - 1
        dims = make_dims(B, int(offsets[2][-1]), int(p.shape[0]), 1, 1, dt, key, offsets=offsets)
    lib = _lib.load()
    nbytes = lib.aknn_index_size(ctypes.byref(dims))
    wbytes = lib.aknn_index_workspace_size(ctypes.byref(dims))
    if nbytes == 0 or wbytes == 0:
        raise RuntimeError(f"aknn_index_size failed: {lib.aknn_last_error().decode()}")
    dev = p.device
    index = torch.empty(nbytes, dtype=torch.uint8, device=dev)
    ws = torch.empty(wbytes, dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_index_build(ctypes.byref(dims), p.data_ptr(), sigma.data_ptr(), R.data_ptr(),
                                        index.data_ptr(), nbytes, ws.data_ptr(), wbytes, _stream(dev)))
    return index


def knn_aniso(q, p, sigma, R, K, *, key: str = "aniso", tune=None):
    """Neighbour search only -> ``[dists, idxs]`` (same order and bits as the forward)."""
    for t, n in zip((q, p, sigma, R), ("q", "p", "sigma", "R")):
        _check_input(t, n)
    dt = _check_float((q, p, sigma, R), ("q", "p", "sigma", "R"))
    f_stub = torch.empty(q.shape[0], p.shape[1] if p.dim() == 3 else 0, 1, dtype=dt, device=q.device)
    B, Q, P, _ = _check_shapes(q, p, f_stub, sigma, R)
    K = _check_k(K, P)
    dev = q.device
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    idxs = torch.empty(B, Q, K, dtype=torch.int64, device=dev)
    if B * Q == 0:
        return [dists, idxs]
    lib = _lib.load()
    dims = make_dims(B, Q, P, 1, K, dt, key)
    ws_bytes = lib.aknn_forward_workspace_size(ctypes.byref(dims))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    tn = _tune(tune)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_knn(ctypes.byref(dims), _tp(tn), q.data_ptr(), p.data_ptr(),
                                sigma.data_ptr(), R.data_ptr(), dists.data_ptr(), idxs.data_ptr(),
                                ws.data_ptr(), ws_bytes, _stream(dev)))
    return [dists, idxs]


def point_dots(rows, f, p, sigma, R, *, offsets=None):
    """Dot records of a broadcast upstream: ``[n, 16]`` = {p, R, sigma, a} per point row
    with ``a[n] = rows[cloud(n)] . f[n]`` (``rows`` [B, F]; ``f`` [B, P, F] and the point
    tensors of the same batch, or packed ragged rows [sum P_i, ...] with ``offsets`` =
    (q_off, p_off, q_off_host, p_off_host)).  ``a`` has the bits the backward kernels
    form.  For the fourth entry of an ``aknn_cache`` tuple: a caller with a fixed
    read-out and fixed points makes them once."""
    for t, n in zip((rows, f, p, sigma, R), ("rows", "f", "p", "sigma", "R")):
        _check_input(t, n)
    dt = _check_float((rows, f, p, sigma, R), ("rows", "f", "p", "sigma", "R"))
    if offsets is None:
        if f.dim() != 3 or rows.dim() != 2 or tuple(rows.shape) != (f.shape[0], f.shape[2]):
            raise RuntimeError("point_dots: rows must be [B, F] for f [B, P, F]")
        B, P, F = f.shape
        dims = make_dims(B, 1, P, F, 1, dt)
        n = B * P
    else:
        B = int(offsets[2].numel()) - 1
        if f.dim() != 2 or rows.dim() != 2 or tuple(rows.shape) != (B, f.shape[1]):
            raise RuntimeError("point_dots: rows must be [B, F] for packed f [sum P_i, F]")
        F = f.shape[1]
        dims = make_dims(B, int(offsets[2][-1]), int(f.shape[0]), F, 1, dt, offsets=offsets)
        n = int(f.shape[0])
    if p.numel() != 3 * n or sigma.numel() != 3 * n or R.numel() != 9 * n:
        raise RuntimeError("point_dots: p, sigma, R must hold 3, 3 and 9 values per point row")
    out = torch.empty(n, 16, dtype=dt, device=f.device)
    lib = _lib.load()
    with torch.cuda.device(f.device):
        _lib.check(lib.aknn_point_dots(ctypes.byref(dims), rows.data_ptr(), f.data_ptr(), p.data_ptr(),
                                       sigma.data_ptr(), R.data_ptr(), out.data_ptr(), _stream(f.device)))
    return out


# ----------------------------------------------------------------------------- CSR
def build_csr(idxs: torch.Tensor, P: int, key: str = "aniso", order=None):
    """Point-side transpose of ``idxs`` [B,Q,K]: ``(offsets int32 [B*P+1], entries int32
    [B*Q*K])``, each point row's entries ``q*K+k`` ascending.  Deterministic.

    With ``order`` (the cache's query order, int32 [B*Q]): entries number ``pos*K+k`` by
    the query's position in that order, and the result is ``(offsets, entries, order)``;
    the backward passes given it (with the same cache) read the query-side inputs in that
    order."""
    _check_input(idxs, "idxs")
    if idxs.dtype != torch.int64 or idxs.dim() != 3:
        raise RuntimeError("idxs must be an int64 tensor [B, Q, K]")
    B, Q, K = idxs.shape
    dev = idxs.device
    off = torch.empty(B * P + 1, dtype=torch.int32, device=dev)
    ent = torch.empty(max(B * Q * K, 1), dtype=torch.int32, device=dev)
    lib = _lib.load()
    dims = make_dims(B, Q, P, 1, K, torch.float32, key)
    ws_bytes = lib.aknn_csr_workspace_size(ctypes.byref(dims))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        if order is not None:
            _lib.check(lib.aknn_build_csr_ordered(ctypes.byref(dims), idxs.data_ptr(), order.data_ptr(), off.data_ptr(),
                                                  ent.data_ptr(), ws.data_ptr(), ws_bytes, _stream(dev)))
            return off, ent, order
        _lib.check(lib.aknn_build_csr(ctypes.byref(dims), idxs.data_ptr(), off.data_ptr(),
                                      ent.data_ptr(), ws.data_ptr(), ws_bytes, _stream(dev)))
    return off, ent


def _cache_order(cache):
    """The cache's query order (for an ordered CSR), or None."""
    return cache[0] if cache is not None and len(cache) >= 1 and cache[0] is not None else None


def _csr_parts(csr, cache):
    """(offsets, entries, needs bit) of a CSR tuple; an ordered CSR must come with the
    cache whose query order it was built in."""
    if csr is None:
        return None, None, 0
    if len(csr) == 3:
        order = _cache_order(cache)
        if order is None or order.data_ptr() != csr[2].data_ptr():
            raise RuntimeError("a CSR built in a query order needs the cache holding that order")
        return csr[0], csr[1], _lib.CSR_QUERY_ORDER
    return csr[0], csr[1], 0


# ----------------------------------------------------------------------------- backward
def _needs_mask(needs, names):
    if needs is None:
        return sum(_NEED_BITS[n] for n in names)
    m = 0
    for n, want in zip(names, needs):
        if want:
            m |= _NEED_BITS[n]
    return m


_NEED_BITS = {
    "dgf": _lib.NEED_DGRAD_F_OUT,
    "dq": _lib.NEED_DQ,
    "dp": _lib.NEED_DP,
    "df": _lib.NEED_DF,
    "dsigma": _lib.NEED_DSIGMA,
    "dR": _lib.NEED_DR,
}


def _check_range(idx: torch.Tensor, P: int, name: str) -> None:
    """Indices in [0, P): one device reduction and a host synchronisation (skip it with
    ``check=False`` where the indices are known good, e.g. under graph capture)."""
    if idx.numel() == 0:
        return
    lo, hi = torch.aminmax(idx)
    if int(lo) < 0 or int(hi) >= P:
        raise RuntimeError(f"{name} must lie in [0, {P}) (got [{int(lo)}, {int(hi)}])")


def _bcast_of(g, q, f):
    """Rows of a broadcast upstream for the batch of q and f, or None."""
    if not isinstance(g, torch.Tensor) or q.dim() != 3 or f.dim() != 3:
        return None
    return broadcast_rows(g, q.shape[0], q.shape[1], f.shape[2])


def _no_queries(names, mask, B, Q, P, F, dt, dev):
    """Gradients of a batch without query rows: nothing reaches the points (zeros) and
    the query-shaped outputs are empty.  No kernel runs."""
    shapes = {"dgf": (B, Q, F), "dq": (B, Q, 3), "dp": (B, P, 3), "df": (B, P, F), "dsigma": (B, P, 3),
              "dR": (B, P, 3, 3)}
    return [torch.zeros(shapes[n], dtype=dt, device=dev) if mask & _NEED_BITS[n] else None for n in names]


def _bwd_common(q, p, f, sigma, R, k_idxs, grads, gnames):
    for t, n in zip((q, p, f, sigma, R, k_idxs) + tuple(grads), ("q", "p", "f", "sigma", "R", "k_idxs") + tuple(gnames)):
        _check_input(t, n)
    dt = _check_float((q, p, f, sigma, R) + tuple(grads), ("q", "p", "f", "sigma", "R") + tuple(gnames))
    B, Q, P, F = _check_shapes(q, p, f, sigma, R)
    if k_idxs.dtype != torch.int64 or k_idxs.dim() != 3 or tuple(k_idxs.shape[:2]) != (B, Q):
        raise RuntimeError(f"k_idxs must be int64 [B, Q, K] (got {tuple(k_idxs.shape)}, {k_idxs.dtype})")
    K = _check_k(int(k_idxs.shape[2]), P)
    return dt, B, Q, P, F, K


def _check_gw(gw, B, Q, K, dt, name="grad_w_out"):
    if gw is None:
        return None
    _check_input(gw, name)
    if gw.dtype != dt or tuple(gw.shape) != (B, Q, K):
        raise RuntimeError(f"{name} must be {dt} [B, Q, K] = {[B, Q, K]} (got {gw.dtype} {tuple(gw.shape)})")
    return gw.contiguous()


def knn_aggregate_aniso_backward(grad_f_out, grad_w_out, q, p, f, sigma, R, w_out, k_idxs, *,
                                 key: str = "aniso", csr=None, needs=None, cache=None, check: bool = True,
                                 use_grad_w_out: bool = False):
    """First-order gradients -> ``[dq, dp, df, dsigma, dR]``.

    Reference: ``knn_aggregate_aniso_backward`` (csrc/knn_aggregate.cpp:101-132,
    csrc/backward.cu:331-389).  Like the reference, ``grad_w_out`` does not enter
    (csrc/backward.cu:93-94) unless ``use_grad_w_out=True``: then ``grad_w_out`` [B,Q,K]
    adds to every weight's upstream (``aknn_backward_w``).  ``w_out`` is not read --
    weights are recomputed from ``k_idxs``.  ``needs`` (5 bools) skips outputs (returned as ``None``);
    ``csr`` = :func:`build_csr` result, built here when point gradients are needed.
    ``grad_f_out`` may be a broadcast view (:func:`broadcast_rows`).  ``k_idxs`` is
    range-checked unless ``check=False`` (the kernels trust it; the autograd Functions
    pass indices from this library's own forward).
    """
    rows = _bcast_of(grad_f_out, q, f)
    dt, B, Q, P, F, K = _bwd_common(q, p, f, sigma, R, k_idxs, (grad_f_out if rows is None else rows,),
                                    ("grad_f_out",))
    if rows is None and tuple(grad_f_out.shape) != (B, Q, F):
        raise RuntimeError(f"grad_f_out must have shape [B, Q, F] (got {tuple(grad_f_out.shape)})")
    names = ("dq", "dp", "df", "dsigma", "dR")
    mask = _needs_mask(needs, names)
    gw = _check_gw(grad_w_out, B, Q, K, dt) if use_grad_w_out else None
    if B * Q == 0:
        return _no_queries(names, mask, B, Q, P, F, dt, q.device)
    if check:
        _check_range(k_idxs, P, "k_idxs")
    slabs = _feature_slabs(F, dt)
    if slabs is not None:
        if mask & _lib.NEED_POINTS and csr is None:
            csr = build_csr(k_idxs, P, key, order=_cache_order(cache))
        df = torch.empty(B, P, F, dtype=dt, device=q.device) if mask & _lib.NEED_DF else None
        res = None
        for s0, s1 in slabs:
            # the weights' own upstream enters once (with the first column range)
            o = knn_aggregate_aniso_backward(_gcols(grad_f_out, rows, s0, s1), gw if s0 == 0 else None, q, p,
                                             _cols(f, s0, s1), sigma, R, w_out, k_idxs, key=key, csr=csr,
                                             needs=needs, cache=cache, check=False,
                                             use_grad_w_out=gw is not None and s0 == 0)
            if df is not None:
                df[..., s0:s1] = o[2]
            if res is None:
                res = o
            else:
                for i in (0, 1, 3, 4):
                    if res[i] is not None:
                        res[i] += o[i]
        res[2] = df
        return res
    dev = q.device
    outs = {
        "dq": torch.empty(B, Q, 3, dtype=dt, device=dev) if mask & _lib.NEED_DQ else None,
        "dp": torch.empty(B, P, 3, dtype=dt, device=dev) if mask & _lib.NEED_DP else None,
        "df": torch.empty(B, P, F, dtype=dt, device=dev) if mask & _lib.NEED_DF else None,
        "dsigma": torch.empty(B, P, 3, dtype=dt, device=dev) if mask & _lib.NEED_DSIGMA else None,
        "dR": torch.empty(B, P, 3, 3, dtype=dt, device=dev) if mask & _lib.NEED_DR else None,
    }
    if mask & _lib.NEED_POINTS and csr is None:
        csr = build_csr(k_idxs, P, key, order=_cache_order(cache))
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt, key)
    off, ent, ordered = _csr_parts(csr, cache)
    mask |= ordered
    P_ = _lib.ptr
    if rows is not None:
        mask |= _lib.GRAD_BROADCAST
        grad_f_out = rows
    ws_bytes = lib.aknn_backward_workspace_size(ctypes.byref(dims), mask)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
    with torch.cuda.device(dev):
        if gw is not None:
            _lib.check(lib.aknn_backward_w(ctypes.byref(dims), mask, grad_f_out.data_ptr(), gw.data_ptr(),
                                           q.data_ptr(), p.data_ptr(), f.data_ptr(), sigma.data_ptr(),
                                           R.data_ptr(), k_idxs.data_ptr(), P_(off), P_(ent),
                                           _cp(_cache_struct(cache)), P_(ws), ws_bytes, P_(outs["dq"]),
                                           P_(outs["dp"]), P_(outs["df"]), P_(outs["dsigma"]), P_(outs["dR"]),
                                           _stream(dev)))
        else:
            _lib.check(lib.aknn_backward(ctypes.byref(dims), mask, grad_f_out.data_ptr(), q.data_ptr(),
                                         p.data_ptr(), f.data_ptr(), sigma.data_ptr(), R.data_ptr(),
                                         k_idxs.data_ptr(), P_(off), P_(ent), _cp(_cache_struct(cache)),
                                         P_(ws), ws_bytes, P_(outs["dq"]),
                                         P_(outs["dp"]), P_(outs["df"]), P_(outs["dsigma"]),
                                         P_(outs["dR"]), _stream(dev)))
    return [outs[n] for n in names]


def knn_aggregate_aniso_backward_2nd(grad_grad_q, grad_f_out, grad_w_out, q, p, f, sigma, R, w_out,
                                     k_idxs, *, key: str = "aniso", csr=None, needs=None, cache=None,
                                     grad_grad_p=None, grad_grad_f=None, grad_grad_sigma=None,
                                     grad_grad_R=None, check: bool = True, use_grad_w_out: bool = False,
                                     need_dgrad_w_out: bool = False):
    """Second-order gradients -> ``[dgrad_f_out, dq, dp, df, dsigma, dR]`` (and
    ``dgrad_w_out`` [B,Q,K] appended with ``use_grad_w_out=True``).

    Reference: ``knn_aggregate_aniso_backward_2nd`` (csrc/knn_aggregate.cpp:135-172,
    csrc/backward.cu:392-455), the gradients of ``<grad_grad_q, dq>``.  The keyword
    upstreams ``grad_grad_p`` [B,P,3], ``grad_grad_f`` [B,P,F], ``grad_grad_sigma`` [B,P,3]
    and ``grad_grad_R`` [B,P,3,3] add the terms the reference drops
    (cuda_knn_aggregate/__init__.py:271-289); ``grad_grad_q`` may then be ``None``.
    ``needs`` (6 bools) skips outputs (``None``).  ``grad_f_out`` and ``check`` as in
    :func:`knn_aggregate_aniso_backward`.  ``use_grad_w_out=True``: the second order of the
    first order with ``grad_w_out`` (``aknn_backward2_w``); the appended ``dgrad_w_out`` is
    computed when ``need_dgrad_w_out`` (else ``None``).
    """
    ups = {"grad_grad_p": (grad_grad_p, 3), "grad_grad_f": (grad_grad_f, None),
           "grad_grad_sigma": (grad_grad_sigma, 3), "grad_grad_R": (grad_grad_R, (3, 3))}
    full = any(t is not None for t, _ in ups.values())
    if grad_grad_q is None and not full:
        raise RuntimeError("grad_grad_q is required (or one of grad_grad_p / _f / _sigma / _R)")
    rows = None if full else _bcast_of(grad_f_out, q, f)
    if full and isinstance(grad_f_out, torch.Tensor) and _bcast_of(grad_f_out, q, f) is not None:
        grad_f_out = grad_f_out.contiguous()  # the full second order takes a materialised upstream
    given = [(t, n) for n, (t, _) in [("grad_grad_q", (grad_grad_q, 0))] + list(ups.items()) if t is not None]
    dt, B, Q, P, F, K = _bwd_common(q, p, f, sigma, R, k_idxs,
                                    tuple(t for t, _ in given) + (grad_f_out if rows is None else rows,),
                                    tuple(n for _, n in given) + ("grad_f_out",))
    if rows is None and tuple(grad_f_out.shape) != (B, Q, F):
        raise RuntimeError(f"grad_f_out must have shape [B, Q, F] (got {tuple(grad_f_out.shape)})")
    if grad_grad_q is not None and tuple(grad_grad_q.shape) != (B, Q, 3):
        raise RuntimeError(f"grad_grad_q must have shape [B, Q, 3] (got {tuple(grad_grad_q.shape)})")
    for n, (t, tail) in ups.items():
        want = (B, P) + ((F,) if tail is None else ((tail,) if isinstance(tail, int) else tail))
        if t is not None and tuple(t.shape) != want:
            raise RuntimeError(f"{n} must have shape {list(want)} (got {tuple(t.shape)})")
    names = ("dgf", "dq", "dp", "df", "dsigma", "dR")
    mask = _needs_mask(needs, names)
    gw = _check_gw(grad_w_out, B, Q, K, dt) if use_grad_w_out else None
    want_dgw = use_grad_w_out and need_dgrad_w_out
    if B * Q == 0:
        res = _no_queries(names, mask, B, Q, P, F, dt, q.device)
        return res + [torch.zeros(B, Q, K, dtype=dt, device=q.device) if want_dgw else None] if use_grad_w_out \
            else res
    if check:
        _check_range(k_idxs, P, "k_idxs")
    slabs = _feature_slabs(F, dt)
    if slabs is not None:
        if (mask & _lib.NEED_POINTS or (full and mask & _lib.NEED_DQ)) and csr is None:
            csr = build_csr(k_idxs, P, key, order=_cache_order(cache))
        dgf = torch.empty(B, Q, F, dtype=dt, device=q.device) if mask & _lib.NEED_DGRAD_F_OUT else None
        df = torch.empty(B, P, F, dtype=dt, device=q.device) if mask & _lib.NEED_DF else None
        res = None
        for s0, s1 in slabs:
            # the weights' own upstream, and its gradient, enter once (first column range)
            first = s0 == 0
            o = knn_aggregate_aniso_backward_2nd(
                grad_grad_q, _gcols(grad_f_out, rows, s0, s1), gw if first else None, q, p, _cols(f, s0, s1), sigma,
                R, w_out, k_idxs,
                key=key, csr=csr, needs=needs, cache=cache, grad_grad_p=grad_grad_p,
                grad_grad_f=_cols(grad_grad_f, s0, s1), grad_grad_sigma=grad_grad_sigma, grad_grad_R=grad_grad_R,
                check=False, use_grad_w_out=use_grad_w_out and first, need_dgrad_w_out=want_dgw and first)
            if dgf is not None:
                dgf[..., s0:s1] = o[0]
            if df is not None:
                df[..., s0:s1] = o[3]
            if res is None:
                res = o
            else:
                for i in (1, 2, 4, 5):
                    if res[i] is not None:
                        res[i] += o[i]
        res[0], res[3] = dgf, df
        return res
    dev = q.device
    outs = {
        "dgf": torch.empty(B, Q, F, dtype=dt, device=dev) if mask & _lib.NEED_DGRAD_F_OUT else None,
        "dq": torch.empty(B, Q, 3, dtype=dt, device=dev) if mask & _lib.NEED_DQ else None,
        "dp": torch.empty(B, P, 3, dtype=dt, device=dev) if mask & _lib.NEED_DP else None,
        "df": torch.empty(B, P, F, dtype=dt, device=dev) if mask & _lib.NEED_DF else None,
        "dsigma": torch.empty(B, P, 3, dtype=dt, device=dev) if mask & _lib.NEED_DSIGMA else None,
        "dR": torch.empty(B, P, 3, 3, dtype=dt, device=dev) if mask & _lib.NEED_DR else None,
    }
    if (mask & _lib.NEED_POINTS or (full and mask & _lib.NEED_DQ)) and csr is None:
        csr = build_csr(k_idxs, P, key, order=_cache_order(cache))
    lib = _lib.load()
    dims = make_dims(B, Q, P, F, K, dt, key)
    off, ent, ordered = _csr_parts(csr, cache)
    mask |= ordered
    P_ = _lib.ptr
    if rows is not None:
        mask |= _lib.GRAD_BROADCAST
        grad_f_out = rows
    ws_bytes = lib.aknn_backward_workspace_size(ctypes.byref(dims), mask | (_lib.SECOND_FULL if full else 0))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
    if use_grad_w_out:
        dgw = torch.empty(B, Q, K, dtype=dt, device=dev) if want_dgw else None
        with torch.cuda.device(dev):
            _lib.check(lib.aknn_backward2_w(
                ctypes.byref(dims), mask | (_lib.NEED_DGRAD_W_OUT if want_dgw else 0), P_(grad_grad_q),
                P_(grad_grad_p), P_(grad_grad_f), P_(grad_grad_sigma), P_(grad_grad_R), grad_f_out.data_ptr(),
                P_(gw), q.data_ptr(), p.data_ptr(), f.data_ptr(), sigma.data_ptr(), R.data_ptr(),
                k_idxs.data_ptr(), P_(off), P_(ent), _cp(_cache_struct(cache)), P_(ws), ws_bytes,
                P_(outs["dgf"]), P_(dgw), P_(outs["dq"]), P_(outs["dp"]), P_(outs["df"]), P_(outs["dsigma"]),
                P_(outs["dR"]), _stream(dev)))
        return [outs[n] for n in names] + [dgw]
    if full:
        with torch.cuda.device(dev):
            _lib.check(lib.aknn_backward2_full(
                ctypes.byref(dims), mask, P_(grad_grad_q), P_(grad_grad_p), P_(grad_grad_f), P_(grad_grad_sigma),
                P_(grad_grad_R), grad_f_out.data_ptr(), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                sigma.data_ptr(), R.data_ptr(), k_idxs.data_ptr(), P_(off), P_(ent), _cp(_cache_struct(cache)),
                P_(ws), ws_bytes, P_(outs["dgf"]), P_(outs["dq"]), P_(outs["dp"]), P_(outs["df"]),
                P_(outs["dsigma"]), P_(outs["dR"]), _stream(dev)))
        return [outs[n] for n in names]
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_backward2(ctypes.byref(dims), mask, grad_grad_q.data_ptr(),
                                      grad_f_out.data_ptr(), q.data_ptr(), p.data_ptr(),
                                      f.data_ptr(), sigma.data_ptr(), R.data_ptr(),
                                      k_idxs.data_ptr(), P_(off), P_(ent), _cp(_cache_struct(cache)),
                                      P_(ws), ws_bytes, P_(outs["dgf"]),
                                      P_(outs["dq"]), P_(outs["dp"]), P_(outs["df"]),
                                      P_(outs["dsigma"]), P_(outs["dR"]), _stream(dev)))
    return [outs[n] for n in names]
