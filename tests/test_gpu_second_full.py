"""GPU parity of the full second order: upstream gradients on dq, dp, df, dsigma and dR
(the reference keeps only the one on dq).  Against FP64 double autograd of the plain
formulation on the same neighbours, and an unrestricted gradgradcheck."""
from __future__ import annotations

import pytest
import torch

import cuda_knn_aggregate as ck
from cuda_knn_aggregate import _C, knn_aggregate_aniso, workloads
from oracle import plain

pytestmark = pytest.mark.gpu
NAMES = ("d_grad_f_out", "dq", "dp", "df", "dsigma", "dR")


def setup(dtype, seed, Q=1500, P=1200, K=8, F=12):
    b = workloads.clustered_batch(Q=Q, P=P, K=K, F=F, seed=seed)
    b = workloads.Batch(*(t.to(dtype) for t in (b.q, b.p, b.f, b.sigma, b.R)), K)
    _, idxs = plain.knn_sort(b.q.float(), b.p.float(), b.sigma.float(), b.R.float(), K)
    gen = torch.Generator().manual_seed(seed + 1)
    g = torch.randn(1, Q, F, generator=gen, dtype=dtype)
    ups = {"h": torch.randn(1, Q, 3, generator=gen, dtype=dtype),
           "hp": torch.randn(1, P, 3, generator=gen, dtype=dtype),
           "hf": torch.randn(1, P, F, generator=gen, dtype=dtype),
           "hs": torch.randn(1, P, 3, generator=gen, dtype=dtype) * 0.1,
           "hR": torch.randn(1, P, 3, 3, generator=gen, dtype=dtype) * 0.1}
    return b, idxs, g, ups


@pytest.mark.parametrize("which", ["all", "no_h", "p_only", "f_only", "sigma_R"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_full_second_order_matches_oracle(cuda, dtype, which):
    b, idxs, g, ups = setup(dtype, seed=31)
    keep = {"all": "h hp hf hs hR", "no_h": "hp hf hs hR", "p_only": "hp", "f_only": "hf",
            "sigma_R": "h hs hR"}[which].split()
    u = {k: (v if k in keep else None) for k, v in ups.items()}
    d = b.to(cuda)
    dev = lambda t: None if t is None else t.to(cuda)
    out = _C.knn_aggregate_aniso_backward_2nd(
        dev(u["h"]), g.to(cuda), None, d.q, d.p, d.f, d.sigma, d.R, None, idxs.to(cuda),
        grad_grad_p=dev(u["hp"]), grad_grad_f=dev(u["hf"]), grad_grad_sigma=dev(u["hs"]),
        grad_grad_R=dev(u["hR"]))
    ins = [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]
    dd = lambda t: None if t is None else t.double()
    ref = plain.backward2_full_autograd(dd(u["h"]), dd(u["hp"]), dd(u["hf"]), dd(u["hs"]), dd(u["hR"]),
                                        g.double(), *ins, idxs)
    tol = 1e-5 if dtype == torch.float32 else 1e-10
    for name, x, y in zip(NAMES, out, ref):
        y = torch.zeros_like(x.cpu().double()) if y is None else y
        err = plain.rel_err(x.cpu().double(), y)
        print(f"  {name}: {err:.2e}", flush=True)
        assert err <= tol, name


def test_full_second_order_with_only_grad_grad_q_matches_reference_path(cuda):
    b, idxs, g, ups = setup(torch.float32, seed=32)
    d = b.to(cuda)
    ref = _C.knn_aggregate_aniso_backward_2nd(ups["h"].to(cuda), g.to(cuda), None, d.q, d.p, d.f, d.sigma, d.R,
                                              None, idxs.to(cuda))
    zero = torch.zeros(1, 1200, 3, device=cuda)
    full = _C.knn_aggregate_aniso_backward_2nd(ups["h"].to(cuda), g.to(cuda), None, d.q, d.p, d.f, d.sigma, d.R,
                                               None, idxs.to(cuda), grad_grad_p=zero)
    for x, y in zip(full, ref):
        assert plain.rel_err(x.cpu(), y.cpu()) < 1e-6


def test_unrestricted_gradgradcheck_through_autograd(cuda):
    """K = P (the neighbour set cannot change under perturbation), FP64, every input
    requiring grad: second derivatives of f_out through the autograd Functions."""
    g = torch.Generator().manual_seed(33)
    B, Q, P, F = 1, 5, 4, 3
    q = torch.rand(B, Q, 3, generator=g, dtype=torch.float64)
    p = torch.rand(B, P, 3, generator=g, dtype=torch.float64)
    f = torch.rand(B, P, F, generator=g, dtype=torch.float64)
    sigma = 0.5 + torch.rand(B, P, 3, generator=g, dtype=torch.float64)
    R = workloads.rand_rot(B * P, g).reshape(B, P, 3, 3).double()
    ins = [t.to(cuda).requires_grad_(True) for t in (q, p, f, sigma, R)]
    fn = lambda *t: knn_aggregate_aniso(*t, P)[0]
    assert ck.FULL_SECOND_ORDER
    assert torch.autograd.gradgradcheck(fn, ins, eps=1e-6, atol=1e-5, rtol=1e-4)


def test_reference_second_order_switch(cuda):
    """FULL_SECOND_ORDER = False drops the upstream on dp..dR like the reference."""
    b, idxs, g, ups = setup(torch.float32, seed=34, Q=300, P=200, K=4, F=4)
    d = b.to(cuda)
    params = [t.clone().requires_grad_(True) for t in (d.q, d.p, d.f, d.sigma, d.R)]
    gv = g.to(cuda)

    def second():
        f_out, *_ = knn_aggregate_aniso(*params, 4)
        first = torch.autograd.grad(f_out, params, gv, create_graph=True)
        L = (first[0] ** 2).sum() + (first[1] ** 2).sum()      # upstream on dq and dp
        return torch.autograd.grad(L, params)

    full = second()
    ck.FULL_SECOND_ORDER = False
    try:
        trunc = second()
    finally:
        ck.FULL_SECOND_ORDER = True
    assert plain.rel_err(full[2].cpu(), trunc[2].cpu()) > 1e-3    # the dp upstream matters for df


def test_eikonal_step_runs_the_dq_only_second_order(cuda):
    """The Eikonal loss depends on dq alone: the point outputs' upstream arrives as
    None, so the second order runs the reference-order kernels, not the full ones."""
    from cuda_knn_aggregate import _lib
    b = workloads.clustered_batch(Q=2000, P=1500, K=8, F=8, seed=35).to(cuda)
    params = [t.clone().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    with _lib.trace() as tr:
        f_out, *_ = knn_aggregate_aniso(*params, 8)
        dq = torch.autograd.grad(f_out, params[0], b.v.expand_as(f_out), create_graph=True)[0]
        L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
        torch.autograd.grad(L, params)
    names = {n for n, _ in tr.records}
    assert "point_grad2" in names and not any(n.endswith("_full") for n in names), names
