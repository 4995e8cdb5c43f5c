"""GPU parity of ``KNNPrecomputedAnisotropicAggregate`` (reference
cuda_knn_aggregate/__init__.py:167-213, csrc/forward.cu:186-230): aggregation with caller-given
neighbour indices, its three outputs, first order and the Eikonal second order through
the same Function, against the FP64 oracle on the same indices.
"""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import KNNPrecomputedAnisotropicAggregate, knn_aggregate_aniso, workloads
from oracle import plain

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def f64(b):
    return [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]


def report(name, got, ref, tol=TOL32):
    err = plain.rel_err(got.detach().cpu().double(), ref.detach().cpu().double())
    print(f"  {name}: rel err {err:.3e}", flush=True)
    assert err <= tol, (name, err)


def _random_idxs(B, Q, P, K, seed):
    """Arbitrary neighbour lists (not nearest, unsorted, repeats allowed)."""
    return torch.randint(P, (B, Q, K), generator=torch.Generator().manual_seed(seed))


def test_precomputed_forward_three_outputs(cuda):
    b = workloads.uniform_batch(2, 900, 400, 12, 32, seed=41, sigma_range=(2.0, 40.0))
    idx = _random_idxs(2, 900, 400, 12, 1)
    x = [t.to(cuda) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    out = KNNPrecomputedAnisotropicAggregate.apply(*x, idx.to(cuda), 12)
    assert isinstance(out, tuple) and len(out) == 3
    f_out, w_out, dists = out
    rf, rw, rD = plain.forward_given_idxs(*f64(b), idx)
    report("f_out", f_out, rf)
    report("w_out", w_out, rw)
    # dists: the FP32 key expression at the given neighbours, bit for bit
    ref_keys = torch.stack([torch.gather(plain.aniso_key(b.q[i], b.p[i], b.sigma[i], b.R[i]), 1, idx[i])
                            for i in range(2)])
    assert torch.equal(dists.cpu(), ref_keys)


def test_precomputed_equals_search_path_on_its_neighbours(cuda):
    b = workloads.bench_small(seed=3)
    x = [t.to(cuda) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = knn_aggregate_aniso(*x, b.K)
    pf, pw, pd = KNNPrecomputedAnisotropicAggregate.apply(*x, idxs, b.K)
    assert torch.equal(pf, f_out) and torch.equal(pw, w_out) and torch.equal(pd, dists)


def test_precomputed_first_and_second_order(cuda):
    b = workloads.uniform_batch(2, 700, 300, 8, 32, seed=42, sigma_range=(2.0, 40.0))
    idx = _random_idxs(2, 700, 300, 8, 2)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    ic = idx.to(cuda)
    f_out, w_out, dists = KNNPrecomputedAnisotropicAggregate.apply(*ins, ic, 8)
    (f_out * b.g.to(cuda)).sum().backward()
    ref = plain.backward_closed(b.g.double(), *f64(b), idx)
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref):
        report("grad " + n, x.grad, y)
        x.grad = None
    # Eikonal second order through the same Function
    f_out = KNNPrecomputedAnisotropicAggregate.apply(*ins, ic, 8)[0]
    dq = torch.autograd.grad(f_out, ins[0], b.v.to(cuda).expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    L.backward()

    def sel(q, p, s, R, K, key="aniso"):
        return None, idx

    Lr, gr = plain.eikonal_step(*f64(b), 8, b.v.double(), select=sel)
    assert abs(L.item() - Lr.item()) <= TOL32 * abs(Lr.item())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, gr):
        report("eikonal grad " + n, x.grad, y)


def test_precomputed_rejects_out_of_range_indices(cuda):
    b = workloads.readme(seed=1)
    x = [t.to(cuda) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    bad = _random_idxs(2, 1000, 500, 12, 3)
    bad[1, 7, 3] = 500
    with pytest.raises(RuntimeError, match="must lie in"):
        KNNPrecomputedAnisotropicAggregate.apply(*x, bad.to(cuda), 12)
    bad[1, 7, 3] = -1
    with pytest.raises(RuntimeError, match="must lie in"):
        KNNPrecomputedAnisotropicAggregate.apply(*x, bad.to(cuda), 12)
