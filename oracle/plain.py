"""Plain-PyTorch CPU oracle for ``knn_aggregate_aniso`` -- TEST INFRASTRUCTURE ONLY.

This module restates, in plain eager PyTorch on the CPU, what the reference computes on
the ``knn_aggregate_aniso`` path.  It is the checker for the HIP library: only the test
suite, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
it.  The shipped package never imports it and has no CPU fallback.

Reference anchors (paths relative to the reference repository):

* key and top-K:       csrc/forward.cu:83-97, csrc/mink.cuh:38-121
* aggregation:         csrc/forward.cu:100-109
* first-order grads:   csrc/backward.cu:31-120 (``grad_w_out`` ignored at :93-94)
* second-order grads:  csrc/backward.cu:152-327 (only ``grad_grad_q`` used,
                       cuda_knn_aggregate/__init__.py:260-289)
* zero-initialised outputs: csrc/forward.cu:310-315, csrc/backward.cu:360-364,424-429
* table lookup:        csrc/forward.cu:113-183 (Euclidean 1-NN, then a [P, K] table row)
* isotropic forward:   csrc/forward.cu:5-53

Contract choices that differ from the reference (documented, not accidental):

* Top-K order is the total order on ``(key, index)``: equal keys are kept and listed by
  lowest point index.  The reference's MinK evicts in slot order on exact ties.
* Keys are evaluated in FP32 (or FP64) with every multiply and add rounded separately,
  in the reference's association order, i.e. without FMA contraction.
* Values and gradients are computed in FP64 given the neighbour indices.

Every function here is pinned by tests in ``tests/test_oracle.py`` against independent
checks (closed forms with hand-computed answers, scipy's KD-tree for the Euclidean special
case, a sequential C selection in ``oracle/knn_ref.c``, FP64 autograd, central finite
differences and ``gradgradcheck``).
"""
from __future__ import annotations

import math

import torch

__all__ = [
    "aniso_key",
    "euclid_key",
    "knn_sort",
    "knn_select",
    "gather_points",
    "forward_given_idxs",
    "backward_closed",
    "backward2_closed",
    "backward_autograd",
    "backward2_autograd",
    "plain_aggregate",
    "eikonal_step",
]


# --------------------------------------------------------------------------- keys
def aniso_key(q: torch.Tensor, p: torch.Tensor, sigma: torch.Tensor, R: torch.Tensor) -> torch.Tensor:
    """Anisotropic key for every (query, point) pair: ``[Qc, P]``.

    Follows csrc/forward.cu:85-92 operation by operation, each op rounded on its own:
    ``x=q0-p0``; ``j=((R00*x)+(R01*y))+(R02*z)`` (rows 1 and 2 likewise);
    ``D=(((j*j)*s0)+((k*k)*s1))+((l*l)*s2)``.
    ``q``: [Qc,3]; ``p``,``sigma``: [P,3]; ``R``: [P,3,3] (row-major ``R[n][row][col]``).
    """
    x = q[:, None, 0] - p[None, :, 0]
    y = q[:, None, 1] - p[None, :, 1]
    z = q[:, None, 2] - p[None, :, 2]
    rows = []
    for r in range(3):
        t = R[None, :, r, 0] * x
        t = t + R[None, :, r, 1] * y
        t = t + R[None, :, r, 2] * z
        rows.append(t)
    j, k, l = rows
    D = (j * j) * sigma[None, :, 0]
    D = D + (k * k) * sigma[None, :, 1]
    D = D + (l * l) * sigma[None, :, 2]
    return D


def euclid_key(q: torch.Tensor, p: torch.Tensor) -> torch.Tensor:
    """Euclidean key ``((x*x)+(y*y))+(z*z)``: csrc/forward.cu:33-36 and :154."""
    x = q[:, None, 0] - p[None, :, 0]
    y = q[:, None, 1] - p[None, :, 1]
    z = q[:, None, 2] - p[None, :, 2]
    return (x * x + y * y) + z * z


def _key(q, p, sigma, R, key):
    if key == "aniso":
        return aniso_key(q, p, sigma, R)
    if key == "euclidean":
        return euclid_key(q, p)
    raise ValueError(f"unknown key {key!r}")


def _chunk_rows(P: int, budget_bytes: int = 1 << 28) -> int:
    return max(1, budget_bytes // max(1, 8 * P))


# --------------------------------------------------------------------------- top-K
def knn_sort(q, p, sigma, R, K: int, key: str = "aniso"):
    """Reference-simple top-K: stable sort of each query's keys, keep the first K.

    Stable sorting breaks equal keys by lowest point index, which is the ``(key, index)``
    total order.  Returns ``dists`` [B,Q,K] in the key dtype and ``idxs`` [B,Q,K] int64.
    """
    B, Q, _ = q.shape
    P = p.shape[1]
    if not (1 <= K <= P):
        raise ValueError("K must satisfy 1 <= K <= P")
    dists = torch.empty(B, Q, K, dtype=q.dtype)
    idxs = torch.empty(B, Q, K, dtype=torch.int64)
    step = _chunk_rows(P)
    for b in range(B):
        for s in range(0, Q, step):
            D = _key(q[b, s:s + step], p[b], sigma[b], R[b], key)
            vals, order = torch.sort(D, dim=1, stable=True)
            dists[b, s:s + step] = vals[:, :K]
            idxs[b, s:s + step] = order[:, :K]
    return dists, idxs


def knn_select(q, p, sigma, R, K: int, key: str = "aniso"):
    """Same result as :func:`knn_sort`, by selection instead of a full sort.

    ``topk`` finds each row's K-th smallest key ``t``; every key ``< t`` is kept, then the
    lowest-index keys equal to ``t`` fill the remaining slots; the K survivors (listed by
    ascending index) are stably sorted by key.  Used for larger oracle runs.
    """
    B, Q, _ = q.shape
    P = p.shape[1]
    if not (1 <= K <= P):
        raise ValueError("K must satisfy 1 <= K <= P")
    dists = torch.empty(B, Q, K, dtype=q.dtype)
    idxs = torch.empty(B, Q, K, dtype=torch.int64)
    step = _chunk_rows(P)
    for b in range(B):
        for s in range(0, Q, step):
            D = _key(q[b, s:s + step], p[b], sigma[b], R[b], key)
            t = torch.topk(D, K, dim=1, largest=False, sorted=True).values[:, K - 1:K]
            lt = D < t
            eq = D == t
            need = K - lt.sum(dim=1, keepdim=True)
            keep = lt | (eq & (torch.cumsum(eq.to(torch.int32), dim=1) <= need))
            rows = D.shape[0]
            cand = torch.nonzero(keep)[:, 1].reshape(rows, K)
            cval = torch.gather(D, 1, cand)
            cval, order = torch.sort(cval, dim=1, stable=True)
            dists[b, s:s + step] = cval
            idxs[b, s:s + step] = torch.gather(cand, 1, order)
    return dists, idxs


# --------------------------------------------------------------------------- aggregation
def gather_points(x: torch.Tensor, idxs: torch.Tensor) -> torch.Tensor:
    """``x`` [B,P,...] gathered at ``idxs`` [B,Q,K] -> [B,Q,K,...] (differentiable)."""
    B = x.shape[0]
    bi = torch.arange(B)[:, None, None]
    return x[bi, idxs]


def _geometry(q, p, sigma, R, idxs):
    pn = gather_points(p, idxs)
    sn = gather_points(sigma, idxs)
    Rn = gather_points(R, idxs)
    d = q[:, :, None, :] - pn                       # [B,Q,K,3]
    u = torch.einsum("bqkij,bqkj->bqki", Rn, d)     # u = R d
    D = (sn * u * u).sum(-1)                        # d^T R^T diag(s) R d
    w = torch.exp(-D)
    v = torch.einsum("bqkij,bqki->bqkj", Rn, sn * u)  # v = R^T (s * u)
    return d, u, D, w, v, sn, Rn


def forward_given_idxs(q, p, f, sigma, R, idxs):
    """``f_out = sum_k w_k f[n_k]``, ``w_out``, ``dists`` given the neighbour indices.

    csrc/forward.cu:100-109 (``w = exp(-D)``; example/equation.png: ``exp(-d^T R^T S R d)``).
    Computed in the inputs' dtype (call with FP64 tensors for the oracle).
    """
    _, _, D, w, _, _, _ = _geometry(q, p, sigma, R, idxs)
    fn = gather_points(f, idxs)
    f_out = (w[..., None] * fn).sum(2)
    return f_out, w, D


def _scatter_points(B, P, idxs, val):
    """Sum ``val`` [B,Q,K,...] into [B,P,...] at ``idxs``."""
    tail = val.shape[3:]
    flat = (idxs + torch.arange(B)[:, None, None] * P).reshape(-1)
    out = torch.zeros((B * P,) + tail, dtype=val.dtype)
    out.index_add_(0, flat, val.reshape((-1,) + tail))
    return out.reshape((B, P) + tail)


def backward_closed(g, q, p, f, sigma, R, idxs, gw=None):
    """First-order gradients by closed form: csrc/backward.cu:31-120.

    With ``a = g . f[n]``: ``dq = -2 sum a w v``; ``dp[n] += 2 a w v``;
    ``dsigma[n]_i += -a w u_i^2``; ``dR[n]_ij += -2 a w s_i u_i d_j``; ``df[n] += w g``.
    ``grad_w_out`` does not enter (csrc/backward.cu:93-94) unless given as ``gw``
    [B,Q,K]: then ``a = g . f[n] + gw`` (the weight's whole upstream; ``df`` unchanged).
    """
    B, P = p.shape[0], p.shape[1]
    d, u, D, w, v, sn, Rn = _geometry(q, p, sigma, R, idxs)
    fn = gather_points(f, idxs)
    a = (g[:, :, None, :] * fn).sum(-1)             # [B,Q,K]
    if gw is not None:
        a = a + gw
    aw = a * w
    dq = -2.0 * (aw[..., None] * v).sum(2)
    dp = _scatter_points(B, P, idxs, 2.0 * aw[..., None] * v)
    dsig = _scatter_points(B, P, idxs, -aw[..., None] * u * u)
    dR = _scatter_points(B, P, idxs, -2.0 * aw[..., None, None] * (sn * u)[..., :, None] * d[..., None, :])
    df = _scatter_points(B, P, idxs, w[..., None] * g[:, :, None, :])
    return dq, dp, df, dsig, dR


def backward2_closed(h, g, q, p, f, sigma, R, idxs, gw=None):
    """Second-order gradients of ``<h, dq>``: csrc/backward.cu:152-327.

    With ``s = h . v``, ``M = R^T diag(s) R`` and ``t = 2 s v - M h``:
    ``dg[q] = sum -2 w s f[n]``; ``df[n] += -2 w s g``; ``dq += 2 a w t``;
    ``dp[n] += -2 a w t``; ``dsigma[n]_i += 2 a w u_i (u_i s - h.R_i)``;
    ``dR[n]_ij += -2 a w s_i (h_j u_i + (h.R_i) d_j - 2 s u_i d_j)``.
    Returns ``(d_grad_f_out, dq, dp, df, dsigma, dR)`` like the reference's
    ``knn_aggregate_aniso_backward_2nd`` (csrc/backward.cu:454).  With ``gw`` (the
    first order's ``grad_w_out``): ``a = g . f[n] + gw`` in every ``a w`` term, and
    ``d_grad_w_out = -2 w s`` is appended.
    """
    B, P = p.shape[0], p.shape[1]
    d, u, D, w, v, sn, Rn = _geometry(q, p, sigma, R, idxs)
    fn = gather_points(f, idxs)
    a = (g[:, :, None, :] * fn).sum(-1)
    if gw is not None:
        a = a + gw
    hk = h[:, :, None, :]
    s = (hk * v).sum(-1)                            # [B,Q,K]
    Mh = torch.einsum("bqkij,bqkj->bqki", Rn.transpose(-1, -2),
                      sn * torch.einsum("bqkij,bqkj->bqki", Rn, hk.expand_as(d)))
    t = 2.0 * s[..., None] * v - Mh
    aw = a * w
    dg = (-2.0 * (w * s)[..., None] * fn).sum(2)
    df = _scatter_points(B, P, idxs, -2.0 * (w * s)[..., None] * g[:, :, None, :])
    dq = (2.0 * aw[..., None] * t).sum(2)
    dp = _scatter_points(B, P, idxs, -2.0 * aw[..., None] * t)
    hR = torch.einsum("bqkij,bqkj->bqki", Rn, hk.expand_as(d))   # h . R_i (row i)
    dsig = _scatter_points(B, P, idxs, 2.0 * aw[..., None] * u * (u * s[..., None] - hR))
    inner = (hk[..., None, :] * u[..., :, None]
             + hR[..., :, None] * d[..., None, :]
             - 2.0 * s[..., None, None] * u[..., :, None] * d[..., None, :])
    dR = _scatter_points(B, P, idxs, -2.0 * aw[..., None, None] * sn[..., :, None] * inner)
    if gw is not None:
        return dg, dq, dp, df, dsig, dR, -2.0 * w * s
    return dg, dq, dp, df, dsig, dR


# --------------------------------------------------------------------------- autograd forms
def plain_aggregate(q, p, f, sigma, R, idxs):
    """Differentiable plain formulation given frozen ``idxs`` (returns ``f_out``)."""
    return forward_given_idxs(q, p, f, sigma, R, idxs)[0]


def backward_autograd(g, q, p, f, sigma, R, idxs):
    """First-order gradients by autograd of :func:`plain_aggregate`."""
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    out = plain_aggregate(*ins, idxs)
    return torch.autograd.grad(out, ins, g)


def backward2_autograd(h, g, q, p, f, sigma, R, idxs):
    """Second-order gradients of ``<h, dq>`` by double autograd of :func:`plain_aggregate`."""
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    gg = g.detach().requires_grad_(True)
    out = plain_aggregate(*ins, idxs)
    dq = torch.autograd.grad(out, ins[0], gg, create_graph=True)[0]
    G = (dq * h).sum()
    return torch.autograd.grad(G, [gg] + ins, allow_unused=True)


def backward2_full_autograd(h, hp, hf, hs, hR, g, q, p, f, sigma, R, idxs):
    """Full second order by double autograd of :func:`plain_aggregate`: the gradients of
    ``<h,dq> + <hp,dp> + <hf,df> + <hs,dsigma> + <hR,dR>`` with respect to
    ``(g, q, p, f, sigma, R)`` (any upstream may be ``None``).  The reference keeps only
    the ``h`` term (cuda_knn_aggregate/__init__.py:271-289)."""
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    gg = g.detach().requires_grad_(True)
    out = plain_aggregate(*ins, idxs)
    first = torch.autograd.grad(out, ins, gg, create_graph=True)
    G = sum((d * u).sum() for d, u in zip(first, (h, hp, hf, hs, hR)) if u is not None)
    return torch.autograd.grad(G, [gg] + ins, allow_unused=True)


def backward_w_autograd(g, gw, q, p, f, sigma, R, idxs):
    """First-order gradients of ``<g, f_out> + <gw, w_out>`` by autograd of the plain
    formulation (the grad_w_out path)."""
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    f_out, w, _ = forward_given_idxs(*ins, idxs)
    return torch.autograd.grad((f_out * g).sum() + (w * gw).sum(), ins)


def backward2_w_autograd(ups, g, gw, q, p, f, sigma, R, idxs):
    """Second order of the grad_w_out path by double autograd: the gradients of
    ``sum_i <ups[i], first[i]>`` (``first`` = (dq, dp, df, dsigma, dR) of
    ``<g, f_out> + <gw, w_out>``; entries of ``ups`` may be ``None``) with respect to
    ``(g, gw, q, p, f, sigma, R)``."""
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    gg = g.detach().requires_grad_(True)
    gww = gw.detach().requires_grad_(True)
    f_out, w, _ = forward_given_idxs(*ins, idxs)
    first = torch.autograd.grad((f_out * gg).sum() + (w * gww).sum(), ins, create_graph=True)
    G = sum((d * u).sum() for d, u in zip(first, ups) if u is not None)
    return torch.autograd.grad(G, [gg, gww] + ins, allow_unused=True)


def eikonal_step(q, p, f, sigma, R, K, v, key="aniso", select=knn_select):
    """One fwd + 1st + 2nd-order step of the Eikonal loss on the plain CPU formulation.

    ``L = sum_q (||d(f_out . v)/dq||_2 - 1)^2`` with ``v`` [F] broadcast over queries.
    Returns ``(L, grads wrt (q, p, f, sigma, R))``.  The neighbour search runs in the
    inputs' dtype; the aggregation and gradients run through autograd.
    """
    with torch.no_grad():
        _, idxs = select(q, p, sigma, R, K, key=key)
    ins = [t.detach().requires_grad_(True) for t in (q, p, f, sigma, R)]
    f_out = plain_aggregate(*ins, idxs)
    gv = v.to(f_out.dtype).expand_as(f_out)
    dq = torch.autograd.grad(f_out, ins[0], gv, create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    grads = torch.autograd.grad(L, ins)
    return L, grads


# --------------------------------------------------------------------------- lookup
def nearest_point(q, p):
    """Euclidean 1-NN per query as the reference's lookup scan finds it
    (csrc/forward.cu:135-157): running minimum started at 1e10 with a strict ``<`` over
    ascending point index, so the lowest index at the minimum wins and index 0 is kept
    when no key is below 1e10 (NaN keys never win).  Returns int64 [B, Q]."""
    B, Q, _ = q.shape
    P = p.shape[1]
    out = torch.empty(B, Q, dtype=torch.int64)
    thr = torch.tensor(1e10, dtype=q.dtype)
    step = _chunk_rows(P)
    for b in range(B):
        for s in range(0, Q, step):
            D = euclid_key(q[b, s:s + step], p[b])
            D = torch.where(D < thr, D, torch.full_like(D, float("inf")))
            vals, idx = torch.min(D, dim=1)          # first index of the minimum
            out[b, s:s + step] = torch.where(torch.isinf(vals), torch.zeros_like(idx), idx)
    return out


def lookup_idxs(q, p, knn_table, K):
    """Neighbours of the table lookup (csrc/forward.cu:160-162): row ``nearest_point``
    of the [P, >=K] table, first K entries, table order."""
    return knn_table[nearest_point(q, p)][..., :K]


# --------------------------------------------------------------------------- isotropic
def iso_forward(q, p, f, sigma, K):
    """The reference's isotropic forward (csrc/forward.cu:5-53): K nearest by Euclidean
    key in the (key, index) order, ``w = exp(-(d^2 * sigma[..., 0]))``, ``f_out = sum w f``,
    ``w_out`` [B,Q,1] = sum of the weights.  Keys in the inputs' dtype (FP32 for the
    neighbours and dists), weights and sums in FP64.  Returns f_out, w_out, dists, idxs."""
    B, P = q.shape[0], p.shape[1]
    dists, idxs = knn_sort(q, p, torch.zeros_like(p), torch.zeros(B, P, 3, 3, dtype=p.dtype), K, key="euclidean")
    s0 = gather_points(sigma[..., 0:1], idxs)[..., 0].double()
    w = torch.exp(-(dists.double() * s0))
    fn = gather_points(f.double(), idxs)
    f_out = (w[..., None] * fn).sum(2)
    return f_out, w.sum(-1, keepdim=True), dists, idxs


def rel_err(a: torch.Tensor, b: torch.Tensor) -> float:
    """``max|a-b| / max|b|`` (0 when both are all-zero): the per-output error measure."""
    a = a.double()
    b = b.double()
    den = b.abs().max().item() if b.numel() else 0.0
    num = (a - b).abs().max().item() if b.numel() else 0.0
    if den == 0.0:
        return 0.0 if num == 0.0 else math.inf
    return num / den
