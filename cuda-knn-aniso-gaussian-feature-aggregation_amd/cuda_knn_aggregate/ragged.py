"""Ragged batches: many point clouds of uneven sizes in one call (offset-packed).

All rows of all clouds are concatenated (``q [sum Q_i, 3]``, ``p [sum P_i, 3]``,
``f [sum P_i, F]``, ``sigma [sum P_i, 3]``, ``R [sum P_i, 3, 3]``) and described by int64
offsets (``q_offsets[b] .. q_offsets[b+1]`` are cloud b's query rows).  Every kernel
covers the whole batch in one launch (no per-cloud loop); ``idxs`` are point indices
within their own cloud, as in the dense call.  Semantics per cloud are those of
:func:`cuda_knn_aggregate.knn_aggregate_aniso`.

The :class:`RaggedLayout` holds the host and device copies of the offsets; build it once
(outside any graph capture) and reuse it: the calls then synchronise nothing and
allocate only through torch, so a whole fwd + bwd + 2nd-order step can be captured.
"""
from __future__ import annotations

import ctypes

import torch

from . import _C, _lib

__all__ = ["RaggedLayout", "knn_aggregate_aniso_ragged", "KNNAnisotropicAggregateRagged",
           "KNNAnisotropicAggregateRaggedBackward"]


class RaggedLayout:
    """Offsets of a ragged batch (host and device copies)."""

    def __init__(self, q_offsets, p_offsets, device):
        qh = torch.as_tensor(q_offsets, dtype=torch.int64).detach().cpu().contiguous()
        ph = torch.as_tensor(p_offsets, dtype=torch.int64).detach().cpu().contiguous()
        if qh.dim() != 1 or ph.dim() != 1 or qh.numel() != ph.numel() or qh.numel() < 2:
            raise RuntimeError("q_offsets and p_offsets must be 1-D with the same length B+1 >= 2")
        if int(qh[0]) != 0 or int(ph[0]) != 0:
            raise RuntimeError("offsets must start at 0")
        if bool((qh[1:] < qh[:-1]).any()) or bool((ph[1:] < ph[:-1]).any()):
            raise RuntimeError("offsets must be non-decreasing")
        self.q_off_host, self.p_off_host = qh, ph
        self.device = torch.device(device)
        self.q_off = qh.to(self.device)
        self.p_off = ph.to(self.device)
        self.B = qh.numel() - 1
        self.Q = int(qh[-1])
        self.P = int(ph[-1])
        self.q_sizes = (qh[1:] - qh[:-1]).tolist()
        self.p_sizes = (ph[1:] - ph[:-1]).tolist()
        self.min_points = min(self.p_sizes)

    def dims(self, F, K, dtype, key="aniso"):
        return _C.make_dims(self.B, self.Q, self.P, F, K, dtype, key,
                            offsets=(self.q_off, self.p_off, self.q_off_host, self.p_off_host))

    def new_cache(self, dtype):
        return _C.new_cache(1, self.Q, self.P, dtype, self.device)


def _check(layout, q, p, f, sigma, R):
    for t, n in zip((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R")):
        _C._check_input(t, n)
    dt = _C._check_float((q, p, f, sigma, R), ("q", "p", "f", "sigma", "R"))
    if tuple(q.shape) != (layout.Q, 3):
        raise RuntimeError(f"q must have shape [{layout.Q}, 3] (got {tuple(q.shape)})")
    P = layout.P
    if tuple(p.shape) != (P, 3) or tuple(sigma.shape) != (P, 3) or tuple(R.shape) != (P, 3, 3):
        raise RuntimeError("p, sigma, R must have shapes [P,3], [P,3], [P,3,3] with P = sum of point counts")
    if f.dim() != 2 or f.shape[0] != P or f.shape[1] < 1:
        raise RuntimeError(f"f must have shape [{P}, F]")
    if q.device != layout.device:
        raise RuntimeError(f"tensors must be on the layout's device {layout.device}")
    return dt, f.shape[1]


def _check_k(layout, K):
    K = int(K)
    if K < 1 or K > 64:
        raise RuntimeError(f"K must be in [1, 64] (got {K})")
    if K > layout.min_points:
        raise RuntimeError(f"K ({K}) must not exceed the smallest cloud's point count ({layout.min_points})")
    return K


def forward(layout, q, p, f, sigma, R, K, *, key="aniso", tune=None, cache=None):
    """-> ``[f_out [Q,F], w_out [Q,K], dists [Q,K], idxs [Q,K] (int64, within-cloud)]``."""
    dt, F = _check(layout, q, p, f, sigma, R)
    K = _check_k(layout, K)
    slabs = _C._feature_slabs(F, dt)
    if slabs is not None:  # feature rows wider than one call: column ranges (each searches again)
        f_out = torch.empty(layout.Q, F, dtype=dt, device=q.device)
        for i, (s0, s1) in enumerate(slabs):
            o = forward(layout, q, p, _C._cols(f, s0, s1), sigma, R, K, key=key, tune=tune, cache=cache)
            f_out[:, s0:s1] = o[0]
            if i == 0:
                w_out, dists, idxs = o[1], o[2], o[3]
        return [f_out, w_out, dists, idxs]
    dev = q.device
    f_out = torch.empty(layout.Q, F, dtype=dt, device=dev)
    w_out = torch.empty(layout.Q, K, dtype=dt, device=dev)
    dists = torch.empty(layout.Q, K, dtype=dt, device=dev)
    idxs = torch.empty(layout.Q, K, dtype=torch.int64, device=dev)
    if layout.Q == 0:
        return [f_out, w_out, dists, idxs]
    lib = _lib.load()
    dims = layout.dims(F, K, dt, key)
    ws_bytes = lib.aknn_forward_workspace_size(ctypes.byref(dims))
    if ws_bytes == 0:
        _lib.check(_lib.AKNN_E_LAUNCH)
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev)
    tn = _C._tune(tune)
    cs = _C._cache_struct(cache)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_forward(ctypes.byref(dims), _C._tp(tn), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                                    sigma.data_ptr(), R.data_ptr(), f_out.data_ptr(), w_out.data_ptr(),
                                    dists.data_ptr(), idxs.data_ptr(), ws.data_ptr(), ws_bytes, _C._cp(cs),
                                    _C._stream(dev)))
    return [f_out, w_out, dists, idxs]


def build_csr(layout, idxs, K, order=None):
    """CSR of a ragged batch (``_C.build_csr``; with ``order`` in that query order)."""
    dev = idxs.device
    off = torch.empty(layout.P + 1, dtype=torch.int32, device=dev)
    ent = torch.empty(max(layout.Q * K, 1), dtype=torch.int32, device=dev)
    lib = _lib.load()
    dims = layout.dims(1, K, torch.float32)
    ws_bytes = lib.aknn_csr_workspace_size(ctypes.byref(dims))
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=dev)
    with torch.cuda.device(dev):
        if order is not None:
            _lib.check(lib.aknn_build_csr_ordered(ctypes.byref(dims), idxs.data_ptr(), order.data_ptr(),
                                                  off.data_ptr(), ent.data_ptr(), ws.data_ptr(), ws_bytes,
                                                  _C._stream(dev)))
            return off, ent, order
        _lib.check(lib.aknn_build_csr(ctypes.byref(dims), idxs.data_ptr(), off.data_ptr(), ent.data_ptr(),
                                      ws.data_ptr(), ws_bytes, _C._stream(dev)))
    return off, ent


def _rows(layout, g):
    """[B, F] rows of a ragged upstream [Q, F] that repeats one row for every query (a
    stride-0 view such as ``v.expand(Q, F)``), or None."""
    if g.dim() != 2 or g.shape[0] < 1 or g.stride(0) != 0 or (g.stride(1) != 1 and g.shape[1] > 1):
        return None
    return g[:1].expand(layout.B, g.shape[1]).contiguous()


def _upstream(g):
    return g if g.dim() == 2 and g.shape[0] >= 1 and g.stride(0) == 0 and (g.stride(1) == 1 or g.shape[1] == 1) \
        else g.contiguous()


def _gcols(g, s0, s1):
    return g[:, s0:s1] if g.stride(0) == 0 else _C._cols(g, s0, s1)


def _backward_impl(layout, second, gg, g, q, p, f, sigma, R, idxs, key, csr, needs, cache, ups=None):
    dt, F = _check(layout, q, p, f, sigma, R)
    K = int(idxs.shape[1])
    names = ("dgf", "dq", "dp", "df", "dsigma", "dR") if second else ("dq", "dp", "df", "dsigma", "dR")
    mask = _C._needs_mask(needs, names)
    full = ups is not None and any(u is not None for u in ups)
    if full and g.stride(0) == 0:
        g = g.contiguous()  # the full second order takes a materialised upstream
    dev = q.device
    Q, P = layout.Q, layout.P
    shapes = {"dgf": (Q, F), "dq": (Q, 3), "dp": (P, 3), "df": (P, F), "dsigma": (P, 3), "dR": (P, 3, 3)}
    if Q == 0:  # no query rows: nothing reaches the points
        return [torch.zeros(shapes[n], dtype=dt, device=dev) if mask & _C._NEED_BITS[n] else None for n in names]
    need_csr = bool(mask & _lib.NEED_POINTS) or (full and bool(mask & _lib.NEED_DQ))
    slabs = _C._feature_slabs(F, dt)
    if slabs is not None:  # column ranges: column outputs side by side, the rest summed in order
        if need_csr and csr is None:
            csr = build_csr(layout, idxs, K, order=_C._cache_order(cache))
        cols = {n: (torch.empty(layout.Q if n == "dgf" else layout.P, F, dtype=dt, device=q.device)
                    if mask & _C._NEED_BITS[n] else None) for n in ("dgf", "df") if n in names}
        res = None
        for s0, s1 in slabs:
            ups_c = None if ups is None else (ups[0], _C._cols(ups[1], s0, s1), ups[2], ups[3])
            o = _backward_impl(layout, second, gg, _gcols(g, s0, s1), q, p, _C._cols(f, s0, s1), sigma, R, idxs,
                               key, csr, needs, cache, ups_c)
            for n, t in cols.items():
                if t is not None:
                    t[:, s0:s1] = o[names.index(n)]
            if res is None:
                res = o
            else:
                for i, n in enumerate(names):
                    if n not in cols and res[i] is not None:
                        res[i] += o[i]
        for n, t in cols.items():
            res[names.index(n)] = t
        return res
    outs = {n: (torch.empty(shapes[n], dtype=dt, device=dev) if mask & _C._NEED_BITS[n] else None) for n in names}
    if need_csr and csr is None:
        csr = build_csr(layout, idxs, K, order=_C._cache_order(cache))
    lib = _lib.load()
    dims = layout.dims(F, K, dt, key)
    off, ent, ordered = _C._csr_parts(csr, cache)
    mask |= ordered
    P_ = _lib.ptr
    rows = None if full else _rows(layout, g)
    if rows is not None:
        mask |= _lib.GRAD_BROADCAST
        g = rows
    ws_bytes = lib.aknn_backward_workspace_size(ctypes.byref(dims), mask | (_lib.SECOND_FULL if full else 0))
    ws = torch.empty(ws_bytes, dtype=torch.uint8, device=dev) if ws_bytes else None
    cs = _C._cp(_C._cache_struct(cache))
    with torch.cuda.device(dev):
        if full:
            rc = lib.aknn_backward2_full(
                ctypes.byref(dims), mask, P_(gg), P_(ups[0]), P_(ups[1]), P_(ups[2]), P_(ups[3]), g.data_ptr(),
                q.data_ptr(), p.data_ptr(), f.data_ptr(), sigma.data_ptr(), R.data_ptr(), idxs.data_ptr(), P_(off),
                P_(ent), cs, P_(ws), ws_bytes, P_(outs["dgf"]), P_(outs["dq"]), P_(outs["dp"]), P_(outs["df"]),
                P_(outs["dsigma"]), P_(outs["dR"]), _C._stream(dev))
        elif second:
            rc = lib.aknn_backward2(ctypes.byref(dims), mask, gg.data_ptr(), g.data_ptr(), q.data_ptr(),
                                    p.data_ptr(), f.data_ptr(), sigma.data_ptr(), R.data_ptr(), idxs.data_ptr(),
                                    P_(off), P_(ent), cs, P_(ws), ws_bytes, P_(outs["dgf"]), P_(outs["dq"]),
                                    P_(outs["dp"]), P_(outs["df"]), P_(outs["dsigma"]), P_(outs["dR"]),
                                    _C._stream(dev))
        else:
            rc = lib.aknn_backward(ctypes.byref(dims), mask, g.data_ptr(), q.data_ptr(), p.data_ptr(),
                                   f.data_ptr(), sigma.data_ptr(), R.data_ptr(), idxs.data_ptr(), P_(off),
                                   P_(ent), cs, P_(ws), ws_bytes, P_(outs["dq"]), P_(outs["dp"]),
                                   P_(outs["df"]), P_(outs["dsigma"]), P_(outs["dR"]), _C._stream(dev))
        _lib.check(rc)
    return [outs[n] for n in names]


def backward(layout, g, q, p, f, sigma, R, idxs, *, key="aniso", csr=None, needs=None, cache=None):
    """First-order gradients -> ``[dq, dp, df, dsigma, dR]``.  ``g`` [Q, F], or a view
    repeating one row for every query (read as such, same results)."""
    return _backward_impl(layout, False, None, _upstream(g), q, p, f, sigma, R, idxs, key, csr, needs, cache)


def backward2(layout, gg, g, q, p, f, sigma, R, idxs, *, key="aniso", csr=None, needs=None, cache=None,
              gg_p=None, gg_f=None, gg_sigma=None, gg_R=None):
    """Second-order gradients -> ``[d grad_f_out, dq, dp, df, dsigma, dR]`` of
    ``<gg, dq> + <gg_p, dp> + <gg_f, df> + <gg_sigma, dsigma> + <gg_R, dR>`` (each upstream
    optional; ``gg`` alone is the reference's second order)."""
    ups = tuple(None if u is None else u.contiguous() for u in (gg_p, gg_f, gg_sigma, gg_R))
    if gg is None and all(u is None for u in ups):
        raise RuntimeError("backward2 needs gg or one of gg_p / gg_f / gg_sigma / gg_R")
    return _backward_impl(layout, True, None if gg is None else gg.contiguous(), _upstream(g), q, p, f, sigma, R,
                          idxs, key, csr, needs, cache, ups)


class KNNAnisotropicAggregateRagged(torch.autograd.Function):
    """Ragged counterpart of :class:`cuda_knn_aggregate.KNNAnisotropicAggregate`."""

    @staticmethod
    def forward(ctx, q, p, f, sigma, R, layout, K, key="aniso", tune=None):
        # (none without query rows: the forward then writes nothing into it)
        cache = layout.new_cache(q.dtype) if any(ctx.needs_input_grad[:5]) and layout.Q > 0 else None
        f_out, w_out, dists, idxs = forward(layout, q, p, f, sigma, R, K, key=key, tune=tune, cache=cache)
        ctx.save_for_backward(q, p, f, sigma, R, idxs, *(cache or ()))
        ctx.layout, ctx.key, ctx.has_cache = layout, key, cache is not None
        ctx.mark_non_differentiable(w_out, dists, idxs)
        ctx.set_materialize_grads(False)
        return f_out, w_out, dists, idxs

    @staticmethod
    def backward(ctx, grad_f_out, grad_w_out, grad_dists, grad_idxs):
        q, p, f, sigma, R, idxs = ctx.saved_tensors[:6]
        cache = tuple(ctx.saved_tensors[6:9]) if ctx.has_cache else None
        if grad_f_out is None:
            return (None,) * 9
        needs = tuple(ctx.needs_input_grad[:5])
        grads = KNNAnisotropicAggregateRaggedBackward.apply(_upstream(grad_f_out), q, p, f, sigma, R, idxs,
                                                            ctx.layout, ctx.key, needs, cache)
        return tuple(grads) + (None,) * 4


class KNNAnisotropicAggregateRaggedBackward(torch.autograd.Function):
    """First order as a differentiable op; its backward is the second order.  As in the
    dense Function, upstream gradients on dp, df, dsigma and dR enter too unless
    ``cuda_knn_aggregate.FULL_SECOND_ORDER`` is False (then only the one on dq, as in the
    reference)."""

    @staticmethod
    def forward(ctx, grad_f_out, q, p, f, sigma, R, idxs, layout, key, needs, cache):
        csr = (build_csr(layout, idxs, int(idxs.shape[1]), order=_C._cache_order(cache))
               if any(needs[1:]) and layout.Q > 0 else None)
        out = backward(layout, grad_f_out, q, p, f, sigma, R, idxs, key=key, csr=csr, needs=needs, cache=cache)
        saved = [grad_f_out, q, p, f, sigma, R, idxs] + (list(csr) if csr else []) + (list(cache) if cache else [])
        ctx.save_for_backward(*saved)
        ctx.layout, ctx.key, ctx.csr_n, ctx.has_cache = layout, key, 0 if csr is None else len(csr), cache is not None
        # outputs the loss does not use arrive as None, not as zeros that would select the
        # full second order
        ctx.set_materialize_grads(False)
        return tuple(out)

    @staticmethod
    def backward(ctx, gg_q, gg_p, gg_f, gg_sigma, gg_R):
        saved = list(ctx.saved_tensors)
        grad_f_out, q, p, f, sigma, R, idxs = saved[:7]
        rest = saved[7:]
        csr = tuple(rest.pop(0) for _ in range(ctx.csr_n)) if ctx.csr_n else None
        cache = tuple(rest[:3]) if ctx.has_cache else None
        ni = ctx.needs_input_grad
        needs = (ni[0], ni[1], ni[2], ni[3], ni[4], ni[5])
        from . import FULL_SECOND_ORDER  # noqa: PLC0415  (the package's switch, read per call)

        ups = (gg_p, gg_f, gg_sigma, gg_R) if FULL_SECOND_ORDER else (None,) * 4
        full = any(u is not None for u in ups)
        if (gg_q is None and not full) or not any(needs):
            return (None,) * 11
        if csr is None and (any(needs[2:]) or (full and needs[1])) and ctx.layout.Q > 0:
            csr = build_csr(ctx.layout, idxs, int(idxs.shape[1]), order=_C._cache_order(cache))
        out = backward2(ctx.layout, gg_q, grad_f_out, q, p, f, sigma, R, idxs, key=ctx.key, csr=csr,
                        needs=needs, cache=cache, gg_p=ups[0], gg_f=ups[1], gg_sigma=ups[2], gg_R=ups[3])
        return tuple(out) + (None,) * 5


def knn_aggregate_aniso_ragged(q, p, f, sigma, R, layout: RaggedLayout, K: int, *, key: str = "aniso",
                               tune=None):
    """k-NN anisotropic aggregation over a ragged batch of clouds (one launch per kernel).

    Returns ``f_out [sum Q_i, F]``, ``w_out``, ``dists``, ``idxs [sum Q_i, K]`` (indices
    within each query's own cloud); differentiable in q, p, f, sigma, R with the same first-
    and second-order semantics as the dense call (including upstream gradients on the
    first-order dp, df, dsigma and dR under ``FULL_SECOND_ORDER``).
    """
    return KNNAnisotropicAggregateRagged.apply(q, p, f, sigma, R, layout, K, key, tune)
