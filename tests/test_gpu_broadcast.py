"""GPU tests of the broadcast upstream (``AKNN_GRAD_BROADCAST``), the Q = 0 batches and the
ragged full second order.

A broadcast upstream (``v.expand_as(f_out)``, the Eikonal read-out) is read as one row per
cloud instead of a materialised [B, Q, F] copy.  Bars:
* dp, dsigma and dR are bit-identical to those of the materialised upstream (the point
  kernels form ``a = g . f[n]`` once per row in the association of the per-entry dot
  products);
* df (the cloud's row times the sum of the entries' coefficients) and dq (summed per query
  from the point dot products) are within 1e-6 of the materialised result and within
  1e-5 of the FP64 oracle;
* d grad_f_out (second order) is bit-identical.
"""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, knn_aggregate_aniso, workloads
from cuda_knn_aggregate import ragged as rg
from oracle import plain

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-10


def f64(b):
    return [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]


def rel(got, ref):
    return plain.rel_err(got.detach().cpu().double(), ref.detach().cpu().double())


def _views(b, dev, per_cloud):
    """(broadcast view, its materialised copy, the [B, F] rows) of a read-out upstream."""
    B, Q, F = b.q.shape[0], b.q.shape[1], b.f.shape[2]
    gen = torch.Generator().manual_seed(5)
    if per_cloud:
        rows = torch.randn(B, 1, F, generator=gen).to(dev, b.q.dtype)
        view = rows.expand(B, Q, F)
    else:
        view = torch.randn(F, generator=gen).to(dev, b.q.dtype).expand(B, Q, F)
    return view, view.contiguous()


NEEDS1 = {"all": None, "dq": (True, False, False, False, False), "points": (False, True, True, True, True),
          "dq+df": (True, False, True, False, False)}


@pytest.mark.parametrize("per_cloud", [False, True])
@pytest.mark.parametrize("F", [16, 64, 100])
@pytest.mark.parametrize("needs", sorted(NEEDS1))
def test_broadcast_first_order(cuda, needs, F, per_cloud):
    b = workloads.uniform_batch(2, 700, 400, 16, F, seed=F, sigma_range=(2.0, 40.0))
    d = b.to(cuda)
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, b.K)
    view, mat = _views(b, cuda, per_cloud)
    assert _C.broadcast_rows(view, *view.shape) is not None and _C.broadcast_rows(mat, *mat.shape) is None
    args = (w_out, d.q, d.p, d.f, d.sigma, d.R, w_out, idxs)
    got = _C.knn_aggregate_aniso_backward(view, *args, needs=NEEDS1[needs])
    ref = _C.knn_aggregate_aniso_backward(mat, *args, needs=NEEDS1[needs])
    ora = plain.backward_closed(mat.cpu().double(), *f64(b), idxs.cpu())
    for n, x, y, z in zip(("dq", "dp", "df", "dsigma", "dR"), got, ref, ora):
        assert (x is None) == (y is None), n
        if x is None:
            continue
        e_mat, e_ora = rel(x, y), rel(x, z)
        print(f"  {n}: bits equal {torch.equal(x, y)}, vs materialised {e_mat:.2e}, vs oracle {e_ora:.2e}",
              flush=True)
        if n in ("dq", "df"):
            assert e_mat <= 1e-6 and e_ora <= TOL32, (n, e_mat, e_ora)
        else:
            assert torch.equal(x, y), n


NEEDS2 = {"all": None, "dq": (False, True, False, False, False, False),
          "points": (False, False, True, True, True, True), "dgf+dq": (True, True, False, False, False, False)}


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("needs", sorted(NEEDS2))
def test_broadcast_second_order(cuda, needs, dtype):
    b = workloads.uniform_batch(2, 600, 300, 12, 32, seed=3, sigma_range=(2.0, 40.0)).to("cpu", dtype)
    d = b.to(cuda)
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, b.K)
    view, mat = _views(b, cuda, False)
    h = torch.randn(2, 600, 3, generator=torch.Generator().manual_seed(8)).to(cuda, dtype)
    args = (w_out, d.q, d.p, d.f, d.sigma, d.R, w_out, idxs)
    got = _C.knn_aggregate_aniso_backward_2nd(h, view, *args, needs=NEEDS2[needs])
    ref = _C.knn_aggregate_aniso_backward_2nd(h, mat, *args, needs=NEEDS2[needs])
    ora = plain.backward2_closed(h.cpu().double(), mat.cpu().double(), *f64(b), idxs.cpu())
    tol = TOL32 if dtype == torch.float32 else TOL64
    for n, x, y, z in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), got, ref, ora):
        assert (x is None) == (y is None), n
        if x is None:
            continue
        e_mat, e_ora = rel(x, y), rel(x, z)
        print(f"  {n}: bits equal {torch.equal(x, y)}, vs materialised {e_mat:.2e}, vs oracle {e_ora:.2e}",
              flush=True)
        if n in ("dq", "df"):
            assert e_mat <= (1e-6 if dtype == torch.float32 else 1e-13) and e_ora <= tol, (n, e_mat, e_ora)
        else:
            assert torch.equal(x, y), n


def test_broadcast_eikonal_step_autograd(cuda):
    """The Eikonal step with v.expand_as(f_out) (read as rows) against the same step with
    the upstream materialised, and both against the FP64 plain autograd."""
    b = workloads.uniform_batch(2, 800, 400, 16, 64, seed=21, sigma_range=(2.0, 40.0))
    grads = {}
    for mode in ("view", "materialised"):
        ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
        f_out, _, _, idxs = knn_aggregate_aniso(*ins, b.K)
        up = b.v.to(cuda).expand_as(f_out)
        if mode == "materialised":
            up = up.contiguous()
        dq = torch.autograd.grad(f_out, ins[0], up, create_graph=True)[0]
        L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
        L.backward()
        grads[mode] = [x.grad for x in ins]

    def sel(q, p, s, R, K, key="aniso"):
        return None, idxs.cpu()

    _, gr = plain.eikonal_step(*f64(b), b.K, b.v.double(), select=sel)
    for n, x, y, z in zip(("q", "p", "f", "sigma", "R"), grads["view"], grads["materialised"], gr):
        e_mat, e_ora = rel(x, y), rel(x, z)
        print(f"  grad {n}: vs materialised {e_mat:.2e}, vs oracle {e_ora:.2e}", flush=True)
        assert e_mat <= 1e-5 and e_ora <= TOL32, (n, e_mat, e_ora)


def test_point_dots_and_cached_dots(cuda):
    """aknn_point_dots equals the dot products (to rounding), and dq with the cached dots
    has the bits of dq with the dots formed per call."""
    b = workloads.uniform_batch(2, 500, 300, 8, 64, seed=31, sigma_range=(2.0, 40.0))
    d = b.to(cuda)
    rows = torch.randn(2, 64, generator=torch.Generator().manual_seed(2)).to(cuda)
    a = _C.point_dots(rows, d.f, d.p, d.sigma, d.R)
    ref = (d.f.double() * rows.double()[:, None, :]).sum(-1).reshape(-1)
    assert rel(a[:, 15], ref) <= 1e-6
    geo = torch.cat([d.p.reshape(-1, 3), d.R.reshape(-1, 9), d.sigma.reshape(-1, 3)], 1)
    assert torch.equal(a[:, :15], geo)
    _, w_out, _, idxs = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, b.K)
    view = rows[:, None, :].expand(2, 500, 64)
    needs = (True, False, False, False, False)
    per_call = _C.knn_aggregate_aniso_backward(view, None, d.q, d.p, d.f, d.sigma, d.R, None, idxs, needs=needs)[0]
    cached = _C.knn_aggregate_aniso_backward(view, None, d.q, d.p, d.f, d.sigma, d.R, None, idxs, needs=needs,
                                             cache=(None, None, None, a))[0]
    assert torch.equal(per_call, cached)


def test_broadcast_ragged(cuda):
    """Ragged batches: one read-out row for all clouds, read as rows; dp, dsigma, dR
    bit-identical to the materialised upstream, dq and df within 1e-6."""
    r = workloads.ragged_small(seed=1)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    f_out, w_out, dists, idxs = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K)
    F = d.f.shape[1]
    view = torch.randn(F, generator=torch.Generator().manual_seed(4)).to(cuda).expand(layout.Q, F)
    got = rg.backward(layout, view, d.q, d.p, d.f, d.sigma, d.R, idxs)
    ref = rg.backward(layout, view.contiguous(), d.q, d.p, d.f, d.sigma, d.R, idxs)
    for n, x, y in zip(("dq", "dp", "df", "dsigma", "dR"), got, ref):
        if n in ("dq", "df"):
            assert rel(x, y) <= 1e-6, n
        else:
            assert torch.equal(x, y), n
    h = torch.randn(layout.Q, 3, generator=torch.Generator().manual_seed(6)).to(cuda)
    got2 = rg.backward2(layout, h, view, d.q, d.p, d.f, d.sigma, d.R, idxs)
    ref2 = rg.backward2(layout, h, view.contiguous(), d.q, d.p, d.f, d.sigma, d.R, idxs)
    for n, x, y in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), got2, ref2):
        if n in ("dq", "df"):
            assert rel(x, y) <= 1e-6, n
        else:
            assert torch.equal(x, y), n


# ----------------------------------------------------------------------------- Q = 0
def test_no_query_rows_dense(cuda):
    """A batch without query rows: forward returns empty outputs, and every backward gives
    zero point gradients and an empty dq, with no kernel reading an unfilled cache."""
    b = workloads.uniform_batch(2, 0, 300, 8, 16, seed=2)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = knn_aggregate_aniso(*ins, 8)
    assert f_out.shape == (2, 0, 16) and idxs.shape == (2, 0, 8)
    (f_out * 2.0).sum().backward()
    assert ins[0].grad is None or ins[0].grad.numel() == 0
    for x in ins[1:]:
        assert x.grad is not None and not bool(x.grad.abs().sum() != 0)
    for x in ins:
        x.grad = None
    f_out = knn_aggregate_aniso(*ins, 8)[0]
    dq = torch.autograd.grad(f_out, ins[0], torch.ones(16, device=cuda).expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    L.backward()
    for x in ins[1:]:
        assert x.grad is None or not bool(x.grad.abs().sum() != 0)
    torch.cuda.synchronize()


def test_no_query_rows_ragged(cuda):
    layout = rg.RaggedLayout([0, 0, 0], [0, 200, 500], cuda)
    gen = torch.Generator().manual_seed(3)
    P, F = 500, 16
    q = torch.zeros(0, 3)
    p = torch.rand(P, 3, generator=gen)
    f = torch.rand(P, F, generator=gen)
    s = 1.0 + torch.rand(P, 3, generator=gen)
    R = workloads.rand_rot(P, gen)
    ins = [t.to(cuda).requires_grad_(True) for t in (q, p, f, s, R)]
    f_out = rg.knn_aggregate_aniso_ragged(*ins, layout, 8)[0]
    assert f_out.shape == (0, F)
    (f_out * 3.0).sum().backward()
    for x in ins[1:]:
        assert x.grad is not None and not bool(x.grad.abs().sum() != 0)
    torch.cuda.synchronize()


# ----------------------------------------------------------------------------- ragged full 2nd order
def test_ragged_full_second_order_matches_dense(cuda):
    """Upstream gradients on the ragged first-order dp, df, dsigma and dR enter the second
    order (as in the dense Function): ragged results equal the dense call per cloud."""
    r = workloads.ragged_small(seed=2)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    _, _, _, idxs = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K)
    gen = torch.Generator().manual_seed(9)
    F = d.f.shape[1]
    g = torch.randn(layout.Q, F, generator=gen).to(cuda)
    h = torch.randn(layout.Q, 3, generator=gen).to(cuda)
    ups = [torch.randn(layout.P, *s, generator=gen).to(cuda) for s in ((3,), (F,), (3,), (3, 3))]
    got = rg.backward2(layout, h, g, d.q, d.p, d.f, d.sigma, d.R, idxs, gg_p=ups[0], gg_f=ups[1], gg_sigma=ups[2],
                       gg_R=ups[3])
    qo, po = r.q_off.tolist(), r.p_off.tolist()
    for c in range(layout.B):
        qs, qe, ps, pe = qo[c], qo[c + 1], po[c], po[c + 1]
        if qe == qs:
            continue
        sl = lambda t, a, e: t[a:e].unsqueeze(0).contiguous()  # noqa: E731
        ref = _C.knn_aggregate_aniso_backward_2nd(
            sl(h, qs, qe), sl(g, qs, qe), None, sl(d.q, qs, qe), sl(d.p, ps, pe), sl(d.f, ps, pe),
            sl(d.sigma, ps, pe), sl(d.R, ps, pe), None, sl(idxs, qs, qe),
            grad_grad_p=sl(ups[0], ps, pe), grad_grad_f=sl(ups[1], ps, pe), grad_grad_sigma=sl(ups[2], ps, pe),
            grad_grad_R=sl(ups[3], ps, pe))
        spans = {"dgf": (qs, qe), "dq": (qs, qe)}
        for n, x, y in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), got, ref):
            a, e = spans.get(n, (ps, pe))
            assert torch.equal(x[a:e], y[0]), (c, n)


def _loss_all(dq, dp, df):
    return dp.square().sum() + df.sin().sum() + dq.square().sum()


def _loss_dp(dq, dp, df):
    return dp.square().sum()  # no upstream on dq: grad_grad_q arrives as None


@pytest.mark.parametrize("loss", [_loss_all, _loss_dp])
def test_ragged_full_second_order_autograd(cuda, loss):
    """A loss on the ragged first-order dp and df reaches the second order through
    autograd (it used to be dropped) and matches the dense autograd per cloud."""
    r = workloads.ragged_small(seed=3)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    ins = [t.to(cuda).requires_grad_(True) for t in (r.q, r.p, r.f, r.sigma, r.R)]
    f_out = rg.knn_aggregate_aniso_ragged(*ins, layout, r.K)[0]
    gen = torch.Generator().manual_seed(12)
    g = torch.randn(f_out.shape, generator=gen).to(cuda)
    dq, dp, df = torch.autograd.grad(f_out, ins[:3], g, create_graph=True)
    loss(dq, dp, df).backward()
    got = [x.grad.clone() for x in ins]
    qo, po = r.q_off.tolist(), r.p_off.tolist()
    for c in range(layout.B):
        qs, qe, ps, pe = qo[c], qo[c + 1], po[c], po[c + 1]
        if qe == qs:
            continue
        cin = [t.detach()[a:e].unsqueeze(0).contiguous().requires_grad_(True)
               for t, (a, e) in zip(ins, [(qs, qe)] + [(ps, pe)] * 4)]
        fo = knn_aggregate_aniso(*cin, r.K)[0]
        cq, cp, cf = torch.autograd.grad(fo, cin[:3], g[qs:qe].unsqueeze(0), create_graph=True)
        loss(cq, cp, cf).backward()
        for n, x, (a, e), y in zip("q p f sigma R".split(), got, [(qs, qe)] + [(ps, pe)] * 4, cin):
            err = rel(x[a:e], y.grad[0])
            assert err <= 1e-6, (c, n, err)
