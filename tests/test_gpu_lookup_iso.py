"""GPU parity of the table-lookup path (knn_aggregate_aniso_lookup) and of the
isotropic forward (_C.knn_aggregate_forward) against the CPU oracle.

Bars: idxs and dists bit-exact (FP32 keys, same expressions); f_out, w_out and the
gradients within 1e-5 of the FP64 oracle on the same neighbours."""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, knn_aggregate_aniso_lookup, workloads
from oracle import plain

pytestmark = pytest.mark.gpu
TOL = 1e-5


def table_of(b, Kt):
    """A [P, Kt] table: each point's Kt nearest points (anisotropic key, oracle)."""
    _, idx = plain.knn_sort(b.p, b.p, b.sigma, b.R, Kt)
    return idx[0].contiguous()


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_lookup_matches_oracle(cuda, dtype):
    b = workloads.clustered_batch(Q=3000, P=2500, K=8, F=16, seed=21)
    b = workloads.Batch(*(t.to(dtype) for t in (b.q, b.p, b.f, b.sigma, b.R)), b.K)
    b.q[0, :5] = torch.tensor([2e5, 0.0, 0.0], dtype=dtype)   # all keys >= 1e10: row 0
    table = table_of(b, 10)
    d = b.to(cuda)
    f_out, w_out, dists, idxs = _C.knn_lookup_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R,
                                                                      table.to(cuda), 8)
    ref_i = plain.lookup_idxs(b.q, b.p, table, 8)
    assert torch.equal(idxs.cpu(), ref_i)
    assert torch.equal(idxs.cpu()[0, :5], table[0, :8].expand(5, 8))
    ins = [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]
    rf, rw, rd = plain.forward_given_idxs(*ins, ref_i)
    if dtype == torch.float32:
        key = plain.aniso_key(b.q[0], b.p[0], b.sigma[0], b.R[0])
        assert torch.equal(dists.cpu()[0], torch.gather(key, 1, ref_i[0]))
    tol = TOL if dtype == torch.float32 else 1e-10
    assert plain.rel_err(f_out.cpu(), rf) <= tol and plain.rel_err(w_out.cpu(), rw) <= tol


def test_lookup_ties_and_duplicates(cuda):
    b = workloads.clustered_batch(Q=800, P=600, K=4, F=4, seed=22)
    for t in (b.p, b.sigma, b.R):
        t[0, 300:400] = t[0, 100:200]                            # duplicated points
    b.q[0, :100] = b.p[0, 300:400]                               # queries exactly on them
    table = table_of(b, 4)
    d = b.to(cuda)
    _, _, _, idxs = _C.knn_lookup_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, table.to(cuda), 4)
    assert torch.equal(idxs.cpu(), plain.lookup_idxs(b.q, b.p, table, 4))


def test_lookup_gradients_first_and_second_order(cuda):
    b = workloads.clustered_batch(Q=1500, P=1200, K=6, F=8, seed=23)
    table = table_of(b, 6)
    d = b.to(cuda)
    params = [t.clone().requires_grad_(True) for t in (d.q, d.p, d.f, d.sigma, d.R)]
    f_out, _, _, idxs = knn_aggregate_aniso_lookup(*params, table.to(cuda), 6)
    v = torch.randn(8, generator=torch.Generator().manual_seed(3)).to(cuda)
    dq = torch.autograd.grad(f_out, params[0], v.expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    grads = torch.autograd.grad(L, params)
    # FP64 plain autograd on the same neighbours
    ins = [t.double().clone().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    fo = plain.plain_aggregate(*ins, idxs.cpu())
    dq_ref = torch.autograd.grad(fo, ins[0], v.cpu().double().expand_as(fo), create_graph=True)[0]
    L_ref = ((dq_ref.norm(dim=-1) - 1.0) ** 2).sum()
    ref = torch.autograd.grad(L_ref, ins)
    assert plain.rel_err(dq.detach().cpu(), dq_ref.detach()) <= TOL
    for name, x, y in zip(("q", "p", "f", "sigma", "R"), grads, ref):
        assert plain.rel_err(x.cpu(), y) <= TOL, name


def test_lookup_errors(cuda):
    b = workloads.clustered_batch(Q=100, P=80, K=4, F=4, seed=24).to(cuda)
    table = torch.zeros(80, 4, dtype=torch.int64, device=cuda)
    with pytest.raises(RuntimeError):                             # K above the table width
        _C.knn_lookup_aggregate_aniso_forward(b.q, b.p, b.f, b.sigma, b.R, table, 5)
    bad = table.clone()
    bad[3, 1] = 80
    with pytest.raises(RuntimeError):                             # entry out of range
        _C.knn_lookup_aggregate_aniso_forward(b.q, b.p, b.f, b.sigma, b.R, bad, 4)


@pytest.mark.parametrize("search", ["brute", "tiles", "wave"])
@pytest.mark.parametrize("K", [1, 5, 16])
def test_iso_forward_matches_oracle(cuda, search, K):
    b = workloads.clustered_batch(Q=2000, P=9000, K=K, F=12, seed=25)
    b.sigma[0, :, 0] = torch.rand(9000, generator=torch.Generator().manual_seed(2)) * 20.0
    d = b.to(cuda)
    f_out, w_out, dists, idxs = _C.knn_aggregate_forward(d.q, d.p, d.f, d.sigma, K, tune={"search": search})
    rf, rw, rd, ri = plain.iso_forward(b.q, b.p, b.f, b.sigma, K)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)
    assert w_out.shape == (1, 2000, 1)
    assert plain.rel_err(f_out.cpu(), rf) <= TOL and plain.rel_err(w_out.cpu(), rw) <= TOL
