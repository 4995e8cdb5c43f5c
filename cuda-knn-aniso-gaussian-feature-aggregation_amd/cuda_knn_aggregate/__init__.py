"""k-NN feature aggregation with anisotropic Gaussian weights on MI355X (gfx950).

Drop-in for the ``knn_aggregate_aniso`` path of the reference package of the same name
(cuda_knn_aggregate/__init__.py:6-32, 64-117, 216-289): same call, same
``torch.autograd.Function`` classes, same first- and second-order backward semantics.
Compute runs in hand-written HIP kernels (``libaknn.so``, C ABI in ``include/aknn.h``)
on the current torch stream; CPU tensors are rejected.

    f_out, w_out, dists, idxs = knn_aggregate_aniso(q, p, f, sigma, R, K)

with ``w = exp(-(q-p)^T R^T diag(sigma) R (q-p))`` over the K neighbours of smallest
anisotropic distance, ordered by (distance, point index).
"""
from __future__ import annotations

import torch

from . import _C, _lib

__all__ = [
    "knn_aggregate_aniso",
    "knn_aggregate_aniso_lookup",
    "KNNAnisotropicAggregate",
    "KNNLookupAnisotropicAggregate",
    "KNNPrecomputedAnisotropicAggregate",
    "KNNAnisotropicAggregateBackward",
]


def knn_aggregate_aniso(
    q: torch.Tensor,
    p: torch.Tensor,
    f: torch.Tensor,
    sigma: torch.Tensor,
    R: torch.Tensor,
    K: int,
    *,
    key: str = "aniso",
    tune=None,
):
    """k-NN feature aggregate using anisotropic gaussian weights.

    Args:
        q: [B, Q, 3] query point coordinates.
        p: [B, P, 3] surface point coordinates.
        f: [B, P, F] feature of surface points.
        sigma: [B, P, 3] anisotropic gaussian scales along all three axes.
        R: [B, P, 3, 3] anisotropic gaussian rotations.
        K: int, number of nearest neighbors (1 <= K <= min(P, 64)).
        key: neighbour ranking, ``"aniso"`` (the reference's, csrc/forward.cu:85-92) or
            ``"euclidean"`` (plain squared distance); weights are anisotropic either way.

    Returns:
        f_out: [B, Q, F] aggregated feature.
        w_out: [B, Q, K] weights. Not differentiable (gradients into it are ignored).
        dists: [B, Q, K] distances. Not differentiable.
        idxs: [B, Q, K] k-NN indices (int64). Not differentiable.
    """
    if isinstance(tune, str):
        if tune != "auto":
            raise ValueError("tune must be None, 'auto' or a dict of search tunables")
        from . import autotune  # noqa: PLC0415

        tune = autotune.tuned(q.detach(), p.detach(), sigma.detach(), R.detach(), K, key=key)
    return KNNAnisotropicAggregate.apply(q, p, f, sigma, R, K, key, tune)


def knn_aggregate_aniso_lookup(
    q: torch.Tensor,
    p: torch.Tensor,
    f: torch.Tensor,
    sigma: torch.Tensor,
    R: torch.Tensor,
    knn_table: torch.Tensor,
    K: int,
):
    """k-NN feature aggregate using anisotropic gaussian weights, with the neighbours
    taken from a pre-computed table (reference: cuda_knn_aggregate/__init__.py:34-62).

    Args:
        q, p, f, sigma, R: as in :func:`knn_aggregate_aniso`.
        knn_table: [P, >=K] int64 table of neighbour indices per surface point (for
            example the kNN of p among itself); shared by the B batches.
        K: number of table entries used per query.

    Each query's Euclidean nearest surface point selects its table row; the first K
    entries of that row are the neighbours (in table order).  Gradients as in
    :func:`knn_aggregate_aniso`, given those neighbours.

    Returns:
        f_out [B, Q, F], w_out [B, Q, K], dists [B, Q, K], idxs [B, Q, K] (int64).
    """
    return KNNLookupAnisotropicAggregate.apply(q, p, f, sigma, R, knn_table, K)


# Second order through KNNAnisotropicAggregateBackward: True honours upstream gradients
# on every first-order output (dq, dp, df, dsigma, dR); False keeps only the one on dq,
# as the reference does (cuda_knn_aggregate/__init__.py:271-289).  The Eikonal loss
# (a function of dq alone) gets the same result either way.
FULL_SECOND_ORDER = True

# Gradients arriving on w_out: False (the reference's behaviour, csrc/backward.cu:93-94
# keeps "+ grad_w_out" commented out) runs the backward without them, so a loss on w_out
# contributes nothing; True adds grad_w_out to each weight's upstream (first order) and
# differentiates that term too (second order, including d grad_w_out).
W_OUT_GRAD = False


def _point_needs(flags):
    return any(flags)


class KNNAnisotropicAggregate(torch.autograd.Function):
    """Neighbour search + aggregation (reference: cuda_knn_aggregate/__init__.py:64-117)."""

    @staticmethod
    def forward(ctx, q, p, f, sigma, R, K, key="aniso", tune=None):
        # The cache (spatial row orders, packed point records) only speeds up the
        # backward kernels; it is filled when a backward can follow.
        # (none without query rows: the forward then writes nothing into it)
        cache = None
        if (any(ctx.needs_input_grad[:5]) and q.is_cuda and p.dim() == 3 and q.dim() == 3
                and q.shape[0] * q.shape[1] > 0):
            cache = _C.new_cache(q.shape[0], q.shape[1], p.shape[1], q.dtype, q.device)
        f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(q, p, f, sigma, R, K, key=key, tune=tune,
                                                                   cache=cache)
        # Like the reference (__init__.py:89); the backward kernels recompute the
        # weights from idxs and never read w_out.
        ctx.save_for_backward(q, p, f, sigma, R, w_out, idxs, *(cache or ()))
        ctx.key = key
        ctx.has_cache = cache is not None
        ctx.set_materialize_grads(False)
        return f_out, w_out, dists, idxs

    @staticmethod
    def backward(ctx, grad_f_out, grad_w_out, grad_dists, grad_idxs):
        q, p, f, sigma, R, w_out, idxs = ctx.saved_tensors[:7]
        cache = tuple(ctx.saved_tensors[7:10]) if getattr(ctx, "has_cache", False) else None
        # Only grad_f_out enters (grad_w_out and grad_dists are ignored, as in the
        # reference, csrc/backward.cu:93-94): no grad on f_out means no gradient at all.
        use_w = W_OUT_GRAD and grad_w_out is not None
        if grad_f_out is None and not use_w:
            return (None,) * 8
        if grad_f_out is None:  # a loss on w_out alone
            grad_f_out = q.new_zeros(()).expand(q.shape[0], q.shape[1], f.shape[2])
        if not use_w:  # never read: a broadcast zero instead of a Q x K fill
            grad_w_out = w_out.new_zeros(()).expand_as(w_out)
        needs = tuple(ctx.needs_input_grad[:5])
        # w_out is detached: the backward never reads it, and an edge from the
        # second-order graph back into this node would only run it again on zeros.
        # A broadcast upstream (v.expand_as(f_out), the Eikonal read-out) stays a view: the
        # library reads its rows instead of a materialised [B, Q, F] copy (same results).
        grad_q, grad_p, grad_f, grad_sigma, grad_R = KNNAnisotropicAggregateBackward.apply(
            _C.upstream(grad_f_out),
            grad_w_out,  # read only under W_OUT_GRAD
            q,
            p,
            f,
            sigma,
            R,
            w_out.detach(),
            idxs,
            ctx.key,
            needs,
            cache,
            use_w,
        )
        return grad_q, grad_p, grad_f, grad_sigma, grad_R, None, None, None


class KNNLookupAnisotropicAggregate(KNNAnisotropicAggregate):
    """Table-lookup neighbours + aggregation (reference: cuda_knn_aggregate/__init__.py:120-165).
    ``apply(q, p, f, sigma, R, knn_table, K) -> (f_out, w_out, dists, idxs)``."""

    @staticmethod
    def forward(ctx, q, p, f, sigma, R, knn_table, K):
        f_out, w_out, dists, idxs = _C.knn_lookup_aggregate_aniso_forward(q, p, f, sigma, R, knn_table, K)
        ctx.save_for_backward(q, p, f, sigma, R, w_out, idxs)
        ctx.key = "aniso"
        ctx.has_cache = False
        ctx.set_materialize_grads(False)
        return f_out, w_out, dists, idxs

    @staticmethod
    def backward(ctx, grad_f_out, grad_w_out, grad_dists, grad_idxs):
        grads = KNNAnisotropicAggregate.backward(ctx, grad_f_out, grad_w_out, grad_dists, grad_idxs)
        return grads[:5] + (None, None)


class KNNPrecomputedAnisotropicAggregate(KNNAnisotropicAggregate):
    """Aggregation with caller-given neighbour indices [B,Q,K] (reference:
    cuda_knn_aggregate/__init__.py:167-213, minus its debug ``torch.save``).
    ``apply(q, p, f, sigma, R, knn_idxs, K) -> (f_out, w_out, dists)``."""

    @staticmethod
    def forward(ctx, q, p, f, sigma, R, knn_idxs, K, key="aniso"):
        f_out, w_out, dists = _C.knn_precomputed_aggregate_aniso_forward(q, p, f, sigma, R, knn_idxs, K, key=key)
        ctx.save_for_backward(q, p, f, sigma, R, w_out, knn_idxs)
        ctx.key = key
        ctx.has_cache = False
        ctx.set_materialize_grads(False)
        return f_out, w_out, dists

    @staticmethod
    def backward(ctx, grad_f_out, grad_w_out, grad_dists):
        grads = KNNAnisotropicAggregate.backward(ctx, grad_f_out, grad_w_out, grad_dists, None)
        return grads[:5] + (None, None, None)


class KNNAnisotropicAggregateBackward(torch.autograd.Function):
    """First-order backward as a differentiable op, whose own backward is the
    second order (reference: cuda_knn_aggregate/__init__.py:216-289).

    The reference honours only ``grad_grad_q`` (__init__.py:260-289); here gradients
    arriving on dp, df, dsigma and dR enter too unless ``FULL_SECOND_ORDER`` is False.
    """

    @staticmethod
    def forward(ctx, grad_f_out, grad_w_out, q, p, f, sigma, R, w_out, idxs, key="aniso",
                needs=None, cache=None, use_w=False):
        if needs is None:
            needs = (True,) * 5
        P = p.shape[1]
        has_q = q.shape[0] * q.shape[1] > 0
        csr = (_C.build_csr(idxs, P, key, order=_C._cache_order(cache)) if _point_needs(needs[1:]) and has_q
               else None)
        dq, dp, df, dsigma, dR = _C.knn_aggregate_aniso_backward(
            grad_f_out, grad_w_out, q, p, f, sigma, R, w_out, idxs, key=key, csr=csr, needs=needs,
            cache=cache, check=False, use_grad_w_out=use_w)
        ctx.use_w = use_w
        saved = [grad_f_out, grad_w_out, q, p, f, sigma, R, w_out, idxs]
        if csr is not None:
            saved += list(csr)
        if cache is not None:
            saved += list(cache)
        ctx.save_for_backward(*saved)
        ctx.csr_n = 0 if csr is None else len(csr)
        ctx.has_cache = cache is not None
        ctx.key = key
        # outputs nothing depends on arrive as None (not zeros): the second order then
        # skips their terms, and the Eikonal loss runs the dq-only kernels
        ctx.set_materialize_grads(False)
        return dq, dp, df, dsigma, dR

    @staticmethod
    def backward(ctx, grad_grad_q, grad_grad_p, grad_grad_f, grad_grad_sigma, grad_grad_R):
        saved = ctx.saved_tensors
        grad_f_out, grad_w_out, q, p, f, sigma, R, w_out, idxs = saved[:9]
        rest = list(saved[9:])
        csr = tuple(rest.pop(0) for _ in range(ctx.csr_n)) if ctx.csr_n else None
        cache = tuple(rest[:3]) if ctx.has_cache else None
        ni = ctx.needs_input_grad
        # outputs: d grad_f_out, dq, dp, df, dsigma, dR (and d grad_w_out on the grad_w_out path)
        needs = (ni[0], ni[2], ni[3], ni[4], ni[5], ni[6])
        need_dgw = ctx.use_w and ni[1]
        ups = (grad_grad_p, grad_grad_f, grad_grad_sigma, grad_grad_R) if FULL_SECOND_ORDER else (None,) * 4
        ups = tuple(None if u is None else u.contiguous() for u in ups)
        full = any(u is not None for u in ups)
        if (grad_grad_q is None and not full) or not (any(needs) or need_dgw):
            return (None,) * 13
        if csr is None and (_point_needs(needs[2:]) or (full and needs[1])) and q.shape[0] * q.shape[1] > 0:
            csr = _C.build_csr(idxs, p.shape[1], ctx.key, order=_C._cache_order(cache))
        res = _C.knn_aggregate_aniso_backward_2nd(
            None if grad_grad_q is None else grad_grad_q.contiguous(),
            _C.upstream(grad_f_out),
            grad_w_out,  # read only on the grad_w_out path
            q,
            p,
            f,
            sigma,
            R,
            w_out,
            idxs,
            key=ctx.key,
            csr=csr,
            needs=needs,
            cache=cache,
            grad_grad_p=ups[0],
            grad_grad_f=ups[1],
            grad_grad_sigma=ups[2],
            grad_grad_R=ups[3],
            check=False,
            use_grad_w_out=ctx.use_w,
            need_dgrad_w_out=need_dgw,
        )
        d_gfo, dq, dp, df, dsigma, dR = res[:6]
        d_gwo = res[6] if ctx.use_w else None
        return d_gfo, d_gwo, dq, dp, df, dsigma, dR, None, None, None, None, None, None
