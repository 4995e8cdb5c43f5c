"""Pins the CPU oracle (oracle/plain.py, oracle/knn_ref.c) against checks that do not
restate it: hand-computed known answers, the matrix form of the README equation,
scipy's KD-tree, an independent sequential selection in C, FP64 autograd, central
finite differences and gradgradcheck.  CPU only."""
from __future__ import annotations

import math

import numpy as np
import pytest
import torch

from oracle import knn_ref, plain
from cuda_knn_aggregate import workloads

torch.set_default_dtype(torch.float32)


def small(B=2, Q=60, P=40, K=6, F=5, seed=0, sigma=(0.05, 0.3)):
    return workloads.uniform_batch(B, Q, P, K, F, seed=seed, sigma_range=sigma)


# ----------------------------------------------------------------------------- key
def test_key_known_answer_row_convention():
    """R[n][row][col] applies as u = R d (csrc/forward.cu:88-90): a non-orthogonal R
    with hand-computed answers; a transposed R or a swapped sigma axis gives other values."""
    R = torch.tensor([[[1.0, 2.0, 0.0], [0.0, 1.0, 0.0], [0.0, 0.0, 1.0]]])
    p = torch.zeros(1, 3)
    q = torch.tensor([[1.0, 1.0, 0.0]])        # d = (1,1,0) -> R d = (3, 1, 0)
    for s, want in (((1.0, 0.0, 0.0), 9.0), ((0.0, 1.0, 0.0), 1.0), ((0.0, 0.0, 1.0), 0.0),
                    ((2.0, 3.0, 5.0), 21.0)):
        D = plain.aniso_key(q, p, torch.tensor([s]), R)
        assert D.item() == want


def test_key_matches_matrix_form_fp64():
    """FP32 key vs d^T R^T diag(s) R d in FP64 numpy (example/equation.png)."""
    b = small(B=1, Q=50, P=70)
    D = plain.aniso_key(b.q[0], b.p[0], b.sigma[0], b.R[0]).double().numpy()
    q, p, s, R = (t[0].double().numpy() for t in (b.q, b.p, b.sigma, b.R))
    d = q[:, None, :] - p[None, :, :]
    M = np.einsum("nji,nj,njk->nik", R, s, R)             # R^T diag(s) R
    ref = np.einsum("qni,nik,qnk->qn", d, M, d)
    assert np.max(np.abs(D - ref) / np.maximum(ref, 1e-12)) < 1e-5


def test_key_no_fused_multiply_add():
    """The FP32 key rounds every product: a case where fma(a,b,c) != round(a*b)+c."""
    a = torch.tensor([1.0 + 2.0 ** -12])
    prod = (a * a).item()              # rounded product
    exact = (1.0 + 2.0 ** -12) ** 2    # needs 25 bits
    assert prod != exact
    # R = a*I, sigma = (1,0,0), q-p = (1,0,0): key = round(round(a*1)^2 * 1)
    R = (a * torch.eye(3))[None]
    D = plain.aniso_key(torch.tensor([[1.0, 0.0, 0.0]]), torch.zeros(1, 3), torch.tensor([[1.0, 0.0, 0.0]]), R)
    assert D.item() == prod


def test_euclidean_special_case():
    """R = I, sigma = 1: the anisotropic key is the Euclidean squared distance."""
    b = small(B=1)
    P = b.p.shape[1]
    I = torch.eye(3).expand(P, 3, 3).contiguous()
    ones = torch.ones(P, 3)
    Da = plain.aniso_key(b.q[0], b.p[0], ones, I)
    De = plain.euclid_key(b.q[0], b.p[0])
    # products by exactly 1 and adds of exact zeros leave the same rounding
    assert torch.equal(Da, De)


# ----------------------------------------------------------------------------- top-K
@pytest.mark.parametrize("key", ["aniso", "euclidean"])
@pytest.mark.parametrize("K", [1, 6, 40])
def test_three_topk_implementations_agree(key, K):
    b = small(B=2, Q=80, P=40)
    d1, i1 = plain.knn_sort(b.q, b.p, b.sigma, b.R, K, key=key)
    d2, i2 = plain.knn_select(b.q, b.p, b.sigma, b.R, K, key=key)
    d3, i3 = knn_ref.knn_ref(b.q, b.p, b.sigma, b.R, K, key=key)
    assert torch.equal(i1, i2) and torch.equal(i1, i3)
    assert torch.equal(d1, d2) and torch.equal(d1, d3)


def test_topk_fp64_agrees():
    b = small().to("cpu", torch.float64)
    d1, i1 = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)
    d3, i3 = knn_ref.knn_ref(b.q, b.p, b.sigma, b.R, 6)
    assert torch.equal(i1, i3) and torch.equal(d1, d3)


def test_euclidean_topk_matches_kdtree():
    """Euclidean neighbours equal scipy's KD-tree on tie-free random points."""
    from scipy.spatial import cKDTree

    b = small(B=1, Q=200, P=300)
    _, idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 8, key="euclidean")
    tree = cKDTree(b.p[0].double().numpy())
    _, ref = tree.query(b.q[0].double().numpy(), k=8)
    assert np.array_equal(idx[0].numpy(), ref)


def test_tie_rule_lowest_index():
    """Equal keys are kept and listed by lowest point index: keys [5,5,1,0], K=3 ->
    (0,1) at 3 and 2, then point 0 (the reference's MinK could keep point 1)."""
    # integer coordinates give exact Euclidean keys 5, 5, 1, 0 around q = 0
    p = torch.tensor([[[2.0, 1.0, 0.0], [1.0, 2.0, 0.0], [1.0, 0.0, 0.0], [0.0, 0.0, 0.0]]])
    q = torch.zeros(1, 1, 3)
    s = torch.ones(1, 4, 3)
    R = torch.eye(3).expand(1, 4, 3, 3).contiguous()
    for fn in (plain.knn_sort, plain.knn_select, knn_ref.knn_ref):
        d, i = fn(q, p, s, R, 3, key="euclidean")
        assert i.tolist() == [[[3, 2, 0]]]
        assert d.tolist() == [[[0.0, 1.0, 5.0]]]


def test_duplicated_points_ties():
    b = small(B=1, Q=50, P=30)
    p = b.p.clone()
    s, R = b.sigma.clone(), b.R.clone()
    p[0, 10:20] = p[0, 0:10]           # exact duplicates (same key for every query)
    s[0, 10:20] = s[0, 0:10]
    R[0, 10:20] = R[0, 0:10]
    d1, i1 = plain.knn_sort(b.q, p, s, R, 12)
    d3, i3 = knn_ref.knn_ref(b.q, p, s, R, 12)
    assert torch.equal(i1, i3)
    # a duplicate pair appears as (n, n+10), lower index first
    for row in i1[0].tolist():
        for n in row:
            if n >= 10 and n < 20 and (n - 10) in row:
                assert row.index(n - 10) < row.index(n)


def test_k_equals_p_returns_every_point():
    b = small(B=1, Q=20, P=9)
    _, i = plain.knn_sort(b.q, b.p, b.sigma, b.R, 9)
    assert torch.equal(i[0].sort(dim=1).values, torch.arange(9).expand(20, 9))


def test_k_out_of_range_raises():
    b = small(B=1, Q=5, P=4)
    for fn in (plain.knn_sort, plain.knn_select, knn_ref.knn_ref):
        with pytest.raises(ValueError):
            fn(b.q, b.p, b.sigma, b.R, 5)
        with pytest.raises(ValueError):
            fn(b.q, b.p, b.sigma, b.R, 0)


# ----------------------------------------------------------------------------- aggregation
def test_forward_k1_closed_form():
    b = small(B=1, Q=10, P=7).to("cpu", torch.float64)
    _, idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 1)
    f_out, w, D = plain.forward_given_idxs(b.q, b.p, b.f, b.sigma, b.R, idx)
    for qi in range(10):
        n = idx[0, qi, 0].item()
        d = (b.q[0, qi] - b.p[0, n]).numpy()
        u = b.R[0, n].numpy() @ d
        Dn = float(np.sum(b.sigma[0, n].numpy() * u * u))
        assert abs(D[0, qi, 0].item() - Dn) < 1e-14
        assert np.allclose(f_out[0, qi].numpy(), math.exp(-Dn) * b.f[0, n].numpy(), rtol=1e-14)


def test_forward_query_on_point_has_unit_weight():
    b = small(B=1, Q=5, P=7).to("cpu", torch.float64)
    q = b.p[:, :5].clone()
    _, idx = plain.knn_sort(q, b.p, b.sigma, b.R, 3)
    assert torch.equal(idx[0, :, 0], torch.arange(5))
    _, w, D = plain.forward_given_idxs(q, b.p, b.f, b.sigma, b.R, idx)
    assert torch.all(w[0, :, 0] == 1.0) and torch.all(D[0, :, 0] == 0.0)


def test_forward_isotropic_gaussian():
    """R arbitrary rotation, sigma = s*(1,1,1): weights are exp(-s |q-p|^2)."""
    b = small(B=1, Q=20, P=15).to("cpu", torch.float64)
    s = torch.full_like(b.sigma, 3.5)
    _, idx = plain.knn_sort(b.q, b.p, s, b.R, 4)
    _, w, _ = plain.forward_given_idxs(b.q, b.p, b.f, s, b.R, idx)
    pn = plain.gather_points(b.p, idx)
    ref = torch.exp(-3.5 * ((b.q[:, :, None] - pn) ** 2).sum(-1))
    # R was rounded to float32, so R^T R = I only to ~1e-7
    assert torch.allclose(w, ref, rtol=1e-6, atol=0)


# ----------------------------------------------------------------------------- gradients
def _f64(b):
    return [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]


def test_first_order_closed_form_matches_autograd():
    b = small(sigma=(2.0, 20.0))
    _, idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)
    ins = _f64(b)
    g = torch.randn(2, 60, 5, dtype=torch.float64)
    c = plain.backward_closed(g, *ins, idx)
    a = plain.backward_autograd(g, *ins, idx)
    for x, y in zip(c, a):
        assert plain.rel_err(x, y) < 1e-12


def test_second_order_closed_form_matches_autograd():
    b = small(sigma=(2.0, 20.0))
    _, idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)
    ins = _f64(b)
    g = torch.randn(2, 60, 5, dtype=torch.float64)
    h = torch.randn(2, 60, 3, dtype=torch.float64)
    c = plain.backward2_closed(h, g, *ins, idx)
    a = plain.backward2_autograd(h, g, *ins, idx)
    for x, y in zip(c, a):
        assert plain.rel_err(x, y) < 1e-12


def test_first_order_central_differences():
    b = small(B=1, Q=8, P=6, K=6, F=3, sigma=(1.0, 5.0))
    ins = _f64(b)
    idx = plain.knn_sort(*[b.q, b.p, b.sigma, b.R], 6)[1]   # K = P: the set never changes
    g = torch.randn(1, 8, 3, dtype=torch.float64)
    grads = plain.backward_closed(g, *ins, idx)
    eps = 1e-6
    for which in range(5):
        x = ins[which]
        flat = x.reshape(-1)
        for j in range(0, flat.numel(), max(1, flat.numel() // 7)):
            old = flat[j].item()
            flat[j] = old + eps
            fp = (plain.forward_given_idxs(*ins, idx)[0] * g).sum().item()
            flat[j] = old - eps
            fm = (plain.forward_given_idxs(*ins, idx)[0] * g).sum().item()
            flat[j] = old
            fd = (fp - fm) / (2 * eps)
            an = grads[which].reshape(-1)[j].item()
            assert abs(fd - an) <= 1e-6 * max(1.0, abs(an)), (which, j, fd, an)


def test_dq_path_gradcheck_and_gradgradcheck():
    """FP64 gradcheck/gradgradcheck of (q,p,f,sigma,R) -> d(f_out . v)/dq with K = P."""
    b = small(B=1, Q=5, P=4, K=4, F=3, sigma=(1.0, 4.0))
    ins = [t.double().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 4)[1]
    v = torch.randn(3, dtype=torch.float64)

    def dq_path(q, p, f, s, R):
        out = plain.plain_aggregate(q, p, f, s, R, idx)
        return torch.autograd.grad(out, q, v.expand_as(out), create_graph=True)[0]

    assert torch.autograd.gradcheck(dq_path, ins, eps=1e-6, atol=1e-8)
    assert torch.autograd.gradgradcheck(dq_path, ins, eps=1e-6, atol=1e-7)


def test_translation_invariance():
    """Moving every query and point together leaves f_out unchanged:
    sum_q dq + sum_n dp = 0."""
    b = small(sigma=(1.0, 9.0))
    idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)[1]
    ins = _f64(b)
    g = torch.randn(2, 60, 5, dtype=torch.float64)
    dq, dp, *_ = plain.backward_closed(g, *ins, idx)
    tot = dq.sum(1) + dp.sum(1)
    assert tot.abs().max().item() < 1e-12 * dq.abs().max().item() * 100
    h = torch.randn(2, 60, 3, dtype=torch.float64)
    _, dq2, dp2, *_ = plain.backward2_closed(h, g, *ins, idx)
    assert (dq2.sum(1) + dp2.sum(1)).abs().max().item() < 1e-10 * dq2.abs().max().item()


def test_df_linear_in_g_and_dsigma_closed_form():
    b = small(sigma=(1.0, 9.0))
    idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)[1]
    ins = _f64(b)
    g1 = torch.randn(2, 60, 5, dtype=torch.float64)
    g2 = torch.randn(2, 60, 5, dtype=torch.float64)
    r1 = plain.backward_closed(g1, *ins, idx)
    r2 = plain.backward_closed(g2, *ins, idx)
    r12 = plain.backward_closed(2.0 * g1 - g2, *ins, idx)
    for a, c, d in zip(r1, r2, r12):
        assert plain.rel_err(d, 2.0 * a - c) < 1e-12


def test_eikonal_step_runs_and_matches_finite_difference():
    b = small(B=1, Q=6, P=5, K=5, F=4, sigma=(1.0, 4.0)).to("cpu", torch.float64)
    L, grads = plain.eikonal_step(b.q, b.p, b.f, b.sigma, b.R, 5, b.v, select=plain.knn_sort)
    eps = 1e-6
    x = b.f.reshape(-1)
    for j in (0, 7, 13):
        old = x[j].item()
        x[j] = old + eps
        Lp = plain.eikonal_step(b.q, b.p, b.f, b.sigma, b.R, 5, b.v, select=plain.knn_sort)[0].item()
        x[j] = old - eps
        Lm = plain.eikonal_step(b.q, b.p, b.f, b.sigma, b.R, 5, b.v, select=plain.knn_sort)[0].item()
        x[j] = old
        assert abs((Lp - Lm) / (2 * eps) - grads[2].reshape(-1)[j].item()) < 1e-6


# ----------------------------------------------------------------------------- lookup
def _nearest_loop(q, p):
    """The reference's scan written out literally (csrc/forward.cu:135-157)."""
    B, Q, P = q.shape[0], q.shape[1], p.shape[1]
    out = torch.zeros(B, Q, dtype=torch.int64)
    for b in range(B):
        for i in range(Q):
            best, arg = torch.tensor(1e10, dtype=q.dtype), 0
            for n in range(P):
                x = q[b, i, 0] - p[b, n, 0]
                y = q[b, i, 1] - p[b, n, 1]
                z = q[b, i, 2] - p[b, n, 2]
                d = (x * x + y * y) + z * z
                if d < best:
                    best, arg = d, n
            out[b, i] = arg
    return out


def test_nearest_point_matches_literal_scan_and_kdtree():
    from scipy.spatial import cKDTree
    b = small(B=2, Q=40, P=30, seed=3)
    q, p = b.q.clone(), b.p.clone()
    p[0, 20] = p[0, 5]                     # duplicate: lowest index must win
    q[0, 0] = p[0, 20]                     # exactly on the duplicated point
    q[1, 3] = torch.tensor([1e6, 0.0, 0.0])  # every key >= 1e10: index 0
    p[1, 7] = float("nan")                 # a NaN point never wins
    got = plain.nearest_point(q, p)
    assert torch.equal(got, _nearest_loop(q, p))
    assert got[0, 0] == 5 and got[1, 3] == 0
    for bi in range(2):                    # distinct distances: the textbook 1-NN
        ok = torch.isfinite(p[bi]).all(-1)
        tree = cKDTree(p[bi][ok].double().numpy())
        _, j = tree.query(q[bi].double().numpy(), k=1)
        ref = torch.nonzero(ok)[:, 0][torch.from_numpy(j)]
        near = (q[bi] - p[bi][got[bi]]).norm(dim=-1) < 1e4
        dg = (q[bi].double() - p[bi][got[bi]].double()).norm(dim=-1)
        dr = (q[bi].double() - p[bi][ref].double()).norm(dim=-1)
        assert torch.equal(dg[near], dr[near])      # same nearest distance everywhere
        uniq = near & (got[bi] != 5) & (got[bi] != 20)  # the tree breaks the planted tie arbitrarily
        assert torch.equal(got[bi][uniq], ref[uniq])


def test_lookup_idxs_takes_the_nearest_points_table_row():
    b = small(B=1, Q=50, P=30, seed=4)
    table = torch.stack([torch.roll(torch.arange(30), -n) for n in range(30)])  # row n starts at n
    idxs = plain.lookup_idxs(b.q, b.p, table, 5)
    n1 = plain.nearest_point(b.q, b.p)
    assert torch.equal(idxs[..., 0], n1)
    assert torch.equal(idxs[..., 1], (n1 + 1) % 30)


# ----------------------------------------------------------------------------- isotropic
def test_iso_forward_k1_closed_form():
    b = small(B=1, Q=20, P=15, seed=5)
    f_out, w_out, dists, idxs = plain.iso_forward(b.q, b.p, b.f, b.sigma, 1)
    n = idxs[0, :, 0]
    d2 = ((b.q[0].double() - b.p[0, n].double()) ** 2).sum(-1)
    w = torch.exp(-d2 * b.sigma[0, n, 0].double())
    assert torch.allclose(w_out[0, :, 0], w, rtol=1e-6, atol=0)
    assert torch.allclose(f_out[0], w[:, None] * b.f[0, n].double(), rtol=1e-6, atol=0)


def test_iso_forward_equals_aniso_with_identity_rotation():
    """R = I and sigma = (s, s, s) turn the anisotropic weight into the isotropic one:
    D = s (x^2 + y^2 + z^2)."""
    b = small(B=2, Q=30, P=25, seed=6)
    s = b.sigma[..., :1].expand(-1, -1, 3).contiguous()
    R = torch.eye(3).expand(2, 25, 3, 3).contiguous()
    f_iso, w_iso, d_iso, i_iso = plain.iso_forward(b.q, b.p, b.f, s, 4)
    _, i_an = plain.knn_sort(b.q, b.p, s, R, 4, key="euclidean")
    f_an, w_an, _ = plain.forward_given_idxs(*(t.double() for t in (b.q, b.p, b.f, s, R)), i_an)
    assert torch.equal(i_iso, i_an)
    assert plain.rel_err(f_iso, f_an) < 1e-6 and plain.rel_err(w_iso[..., 0], w_an.sum(-1)) < 1e-6


# ----------------------------------------------------------------------------- full second order
def test_full_second_order_reduces_to_the_reference_terms():
    """With an upstream on dq only, the full second order is the reference's."""
    b = small(B=1, Q=12, P=9, K=4, F=3, seed=7)
    ins = [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]
    _, idxs = plain.knn_sort(b.q, b.p, b.sigma, b.R, 4)
    g = torch.randn(1, 12, 3, dtype=torch.float64)
    h = torch.randn(1, 12, 3, dtype=torch.float64)
    full = plain.backward2_full_autograd(h, None, None, None, None, g, *ins, idxs)
    ref = plain.backward2_closed(h, g, *ins, idxs)
    for x, y in zip(full, ref):
        assert plain.rel_err(x, y) < 1e-12


def test_full_second_order_gradgradcheck():
    """Second derivatives of the whole map (q, p, f, sigma, R) -> f_out, K = P (no
    neighbour changes under perturbation), checked by finite differences."""
    b = small(B=1, Q=4, P=3, K=3, F=2, seed=8)
    idxs = torch.arange(3).expand(1, 4, 3).contiguous()
    ins = [t.double().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    assert torch.autograd.gradgradcheck(lambda *t: plain.plain_aggregate(*t, idxs), ins, eps=1e-6, atol=1e-5)
    g = torch.randn(1, 4, 2, dtype=torch.float64)
    ups = [torch.randn_like(t) for t in ins]
    got = plain.backward2_full_autograd(*ups, g, *ins, idxs)
    assert all(x is not None for x in got)


# ----------------------------------------------------------------------------- grad_w_out path
def test_w_out_path_closed_forms_match_autograd():
    """With grad_w_out, the closed forms (a = g.f[n] + gw) equal FP64 autograd of
    <g, f_out> + <gw, w_out>, first and second order (d grad_w_out included)."""
    b = small(sigma=(2.0, 20.0), seed=3)
    _, idx = plain.knn_sort(b.q, b.p, b.sigma, b.R, 6)
    ins = _f64(b)
    g = torch.randn(2, 60, 5, dtype=torch.float64)
    gw = torch.randn(2, 60, 6, dtype=torch.float64)
    c = plain.backward_closed(g, *ins, idx, gw=gw)
    a = plain.backward_w_autograd(g, gw, *ins, idx)
    for x, y in zip(c, a):
        assert plain.rel_err(x, y) < 1e-12
    h = torch.randn(2, 60, 3, dtype=torch.float64)
    c2 = plain.backward2_closed(h, g, *ins, idx, gw=gw)      # dg, dq, dp, df, ds, dR, dgw
    a2 = plain.backward2_w_autograd((h, None, None, None, None), g, gw, *ins, idx)  # dg, dgw, dq..dR
    order = [0, 6, 1, 2, 3, 4, 5]
    for i, y in zip(order, a2):
        assert plain.rel_err(c2[i], y) < 1e-12, i
    # gw = 0 gives the reference's gradients
    z = plain.backward_closed(g, *ins, idx, gw=torch.zeros_like(gw))
    for x, y in zip(z, plain.backward_closed(g, *ins, idx)):
        assert plain.rel_err(x, y) < 1e-15


def test_w_out_path_central_differences():
    """d<gw, w_out>/dx by central differences of the FP64 forward."""
    b = small(B=1, Q=8, P=6, K=6, F=3, sigma=(1.0, 5.0), seed=4)
    ins = _f64(b)
    idx = plain.knn_sort(*[b.q, b.p, b.sigma, b.R], 6)[1]
    g = torch.zeros(1, 8, 3, dtype=torch.float64)
    gw = torch.randn(1, 8, 6, dtype=torch.float64)
    grads = plain.backward_closed(g, *ins, idx, gw=gw)
    eps = 1e-6
    for which in (0, 1, 3, 4):
        flat = ins[which].reshape(-1)
        for j in range(0, flat.numel(), max(1, flat.numel() // 5)):
            old = flat[j].item()
            flat[j] = old + eps
            fp = (plain.forward_given_idxs(*ins, idx)[1] * gw).sum().item()
            flat[j] = old - eps
            fm = (plain.forward_given_idxs(*ins, idx)[1] * gw).sum().item()
            flat[j] = old
            an = grads[which].reshape(-1)[j].item()
            assert abs((fp - fm) / (2 * eps) - an) <= 1e-6 * max(1.0, abs(an)), (which, j)
    assert float(grads[2].abs().max()) == 0.0  # w_out does not depend on f
