"""GPU parity of the grad_w_out path (no counterpart in the reference, whose first order
keeps "+ grad_w_out" commented out at csrc/backward.cu:93-94): with
``cuda_knn_aggregate.W_OUT_GRAD = True`` a loss on w_out reaches every gradient through
the weights' upstream, first and second order (d grad_w_out included), checked against
FP64 autograd of the oracle; with the default it contributes nothing, as in the reference.
"""
from __future__ import annotations

import pytest
import torch

import cuda_knn_aggregate as ck
from cuda_knn_aggregate import _C, knn_aggregate_aniso, workloads
from oracle import plain

pytestmark = pytest.mark.gpu

TOL = 1e-5


@pytest.fixture
def w_out_grad():
    old = ck.W_OUT_GRAD
    ck.W_OUT_GRAD = True
    try:
        yield
    finally:
        ck.W_OUT_GRAD = old


def report(name, got, ref, tol=TOL):
    err = plain.rel_err(got.detach().cpu().double(), ref.detach().cpu().double())
    print(f"  {name}: rel err {err:.3e}", flush=True)
    assert err <= tol, (name, err)


def _batch(F=16, seed=51):
    return workloads.uniform_batch(2, 700, 300, 8, F, seed=seed, sigma_range=(2.0, 40.0))


def _f64(b):
    return [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]


def test_default_ignores_w_out(cuda):
    b = _batch()
    gw = torch.randn(2, 700, 8).to(cuda)
    g = torch.randn(2, 700, 16).to(cuda)
    res = []
    for with_w in (False, True):
        ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
        f_out, w_out, _, _ = knn_aggregate_aniso(*ins, 8)
        L = (f_out * g).sum() + ((w_out * gw).sum() if with_w else 0.0)
        L.backward()
        res.append([x.grad for x in ins])
    for x, y in zip(*res):
        assert torch.equal(x, y)


@pytest.mark.parametrize("bcast", [False, True])
def test_first_order(cuda, w_out_grad, bcast):
    b = _batch()
    gen = torch.Generator().manual_seed(3)
    g = (torch.randn(16, generator=gen).expand(2, 700, 16) if bcast else torch.randn(2, 700, 16, generator=gen))
    gw = torch.randn(2, 700, 8, generator=gen)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, _, idxs = knn_aggregate_aniso(*ins, 8)
    gd = g.to(cuda)
    up = gd if not bcast else gd[0, 0].expand_as(f_out)
    torch.autograd.backward([f_out, w_out], [up, gw.to(cuda)])
    ref = plain.backward_w_autograd(g.double().contiguous(), gw.double(), *_f64(b), idxs.cpu())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref):
        report(n, x.grad, y)


@pytest.mark.parametrize("bcast", [False, True])
def test_second_order_dq(cuda, w_out_grad, bcast):
    """<h, dq> of the first order of <g, f_out> + <gw, w_out>: gradients of q, p, f,
    sigma, R and of gw itself (the Eikonal shape when g is broadcast)."""
    b = _batch(seed=52)
    gen = torch.Generator().manual_seed(4)
    g = (torch.randn(16, generator=gen).expand(2, 700, 16) if bcast else torch.randn(2, 700, 16, generator=gen))
    gw = torch.randn(2, 700, 8, generator=gen)
    h = torch.randn(2, 700, 3, generator=gen)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    gwd = gw.to(cuda).requires_grad_(True)
    f_out, w_out, _, idxs = knn_aggregate_aniso(*ins, 8)
    gd = g.to(cuda)
    up = gd if not bcast else gd[0, 0].expand_as(f_out)
    dq = torch.autograd.grad([f_out, w_out], ins[0], [up, gwd], create_graph=True)[0]
    (dq * h.to(cuda)).sum().backward()
    ref = plain.backward2_w_autograd((h.double(), None, None, None, None), g.double().contiguous(), gw.double(),
                                     *_f64(b), idxs.cpu())
    report("d grad_w_out", gwd.grad, ref[1])
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref[2:]):
        report(n, x.grad, y)


def test_full_second_order(cuda, w_out_grad):
    """Upstream on dq and dp (full second order) with grad_w_out."""
    b = _batch(seed=53)
    gen = torch.Generator().manual_seed(5)
    g = torch.randn(2, 700, 16, generator=gen)
    gw = torch.randn(2, 700, 8, generator=gen)
    h = torch.randn(2, 700, 3, generator=gen)
    hp = torch.randn(2, 300, 3, generator=gen)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    gwd = gw.to(cuda).requires_grad_(True)
    f_out, w_out, _, idxs = knn_aggregate_aniso(*ins, 8)
    dq, dp = torch.autograd.grad([f_out, w_out], ins[:2], [g.to(cuda), gwd], create_graph=True)
    ((dq * h.to(cuda)).sum() + (dp * hp.to(cuda)).sum()).backward()
    ref = plain.backward2_w_autograd((h.double(), hp.double(), None, None, None), g.double(), gw.double(),
                                     *_f64(b), idxs.cpu())
    report("d grad_w_out", gwd.grad, ref[1])
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref[2:]):
        report(n, x.grad, y)


def test_wide_features_enter_once(cuda):
    """Feature widths split into column ranges: the weights' upstream is added once."""
    b = workloads.uniform_batch(1, 300, 200, 8, 600, seed=54, sigma_range=(2.0, 40.0))
    x = [t.to(cuda) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    _, _, _, idxs = _C.knn_aggregate_aniso_forward(*x, 8)
    gen = torch.Generator().manual_seed(6)
    g = torch.randn(1, 300, 600, generator=gen)
    gw = torch.randn(1, 300, 8, generator=gen)
    h = torch.randn(1, 300, 3, generator=gen)
    got = _C.knn_aggregate_aniso_backward(g.to(cuda), gw.to(cuda), *x, None, idxs, use_grad_w_out=True)
    ref = plain.backward_closed(g.double(), *_f64(b), idxs.cpu(), gw=gw.double())
    for n, a, r in zip(("dq", "dp", "df", "dsigma", "dR"), got, ref):
        report(n, a, r)
    got2 = _C.knn_aggregate_aniso_backward_2nd(h.to(cuda), g.to(cuda), gw.to(cuda), *x, None, idxs,
                                               use_grad_w_out=True, need_dgrad_w_out=True)
    ref2 = plain.backward2_closed(h.double(), g.double(), *_f64(b), idxs.cpu(), gw=gw.double())
    for n, a, r in zip(("dgf", "dq", "dp", "df", "dsigma", "dR", "dgw"), got2, ref2):
        report(n, a, r)


def test_unrestricted_gradgradcheck(cuda, w_out_grad):
    """With the grad_w_out path and the full second order, torch's gradcheck and
    gradgradcheck pass on (f_out, w_out) as functions of every input (FP64, K = P so the
    neighbour sets do not change under the perturbations)."""
    assert ck.FULL_SECOND_ORDER
    b = workloads.uniform_batch(1, 5, 4, 4, 3, seed=55, sigma_range=(1.0, 5.0))
    ins = [t.double().to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]

    def fn(*x):
        f_out, w_out, _, _ = knn_aggregate_aniso(*x, 4)
        return f_out, w_out

    assert torch.autograd.gradcheck(fn, ins, eps=1e-6, atol=1e-5, rtol=1e-4)
    assert torch.autograd.gradgradcheck(fn, ins, eps=1e-6, atol=1e-5, rtol=1e-4)
