"""GPU tests of ragged batches and graph capture.

A ragged batch must give, row for row, the same bits as calling the dense path once per
cloud (same kernels, same per-row arithmetic, same fixed reduction orders), and must
match the CPU oracle per cloud.  A captured fwd + 1st + 2nd-order step must reproduce the
eager step on every replay.
"""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, knn_aggregate_aniso, workloads
from cuda_knn_aggregate.ragged import RaggedLayout, knn_aggregate_aniso_ragged
from cuda_knn_aggregate import ragged as rg
from oracle import knn_ref, plain

pytestmark = pytest.mark.gpu


def small_ragged(seed=0, with_empty=False):
    r = workloads.ragged_small(seed)
    if with_empty:  # a cloud without queries in the middle
        q_off = r.q_off.clone()
        b = 2
        n = int(q_off[b + 1] - q_off[b])
        keep = torch.cat([torch.arange(0, int(q_off[b])), torch.arange(int(q_off[b + 1]), int(q_off[-1]))])
        r.q = r.q[keep].contiguous()
        q_off[b + 1:] -= n
        r.q_off = q_off
    return r


@pytest.mark.parametrize("search", ["brute", "tiles", "wave"])
@pytest.mark.parametrize("with_empty", [False, True])
def test_ragged_forward_equals_per_cloud(cuda, search, with_empty):
    r = small_ragged(1, with_empty)
    layout = RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    out = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K, tune={"search": search})
    for b in range(r.B):
        c = r.cloud(b)
        if c.q.shape[1] == 0:
            continue
        cd = c.to(cuda)
        ref = _C.knn_aggregate_aniso_forward(cd.q.contiguous(), cd.p.contiguous(), cd.f.contiguous(),
                                             cd.sigma.contiguous(), cd.R.contiguous(), r.K, tune={"search": search})
        qs, qe = int(r.q_off[b]), int(r.q_off[b + 1])
        for x, y in zip(out, ref):
            assert torch.equal(x[qs:qe], y[0]), b
        rd, ri = knn_ref.knn_ref(c.q, c.p, c.sigma, c.R, r.K)
        assert torch.equal(out[3][qs:qe].cpu(), ri[0]) and torch.equal(out[2][qs:qe].cpu(), rd[0])


def test_ragged_backward_equals_per_cloud(cuda):
    r = small_ragged(2)
    layout = RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    F = r.f.shape[1]
    f_out, w_out, dists, idxs = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K)
    g = torch.randn(layout.Q, F, device=cuda)
    h = torch.randn(layout.Q, 3, device=cuda)
    g1 = rg.backward(layout, g, d.q, d.p, d.f, d.sigma, d.R, idxs)
    g2 = rg.backward2(layout, h, g, d.q, d.p, d.f, d.sigma, d.R, idxs)
    for b in range(r.B):
        qs, qe = int(r.q_off[b]), int(r.q_off[b + 1])
        ps, pe = int(r.p_off[b]), int(r.p_off[b + 1])
        c = r.cloud(b)
        ins = [t.double() for t in (c.q, c.p, c.f, c.sigma, c.R)]
        ib = idxs[qs:qe].cpu()[None]
        ref1 = plain.backward_closed(g[qs:qe].cpu().double()[None], *ins, ib)
        for n, x, y, rows in zip(("dq", "dp", "df", "dsigma", "dR"), g1, ref1,
                                 [(qs, qe)] + [(ps, pe)] * 4):
            assert plain.rel_err(x[rows[0]:rows[1]].cpu()[None], y) < 1e-5, (b, n)
        ref2 = plain.backward2_closed(h[qs:qe].cpu().double()[None], g[qs:qe].cpu().double()[None], *ins, ib)
        for n, x, y, rows in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), g2, ref2,
                                 [(qs, qe), (qs, qe)] + [(ps, pe)] * 4):
            assert plain.rel_err(x[rows[0]:rows[1]].cpu()[None], y) < 1e-5, (b, n)


def test_ragged_wide_feature_rows(cuda):
    """Feature rows wider than one library call (F = 600: column ranges of 512 and 88)
    through the ragged forward and both backward orders, vs the FP64 oracle per cloud."""
    r = small_ragged(3)
    F = 600
    r.f = torch.rand(r.f.shape[0], F, generator=torch.Generator().manual_seed(33))
    layout = RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    f_out, w_out, dists, idxs = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K)
    g = torch.randn(layout.Q, F, device=cuda)
    h = torch.randn(layout.Q, 3, device=cuda)
    g1 = rg.backward(layout, g, d.q, d.p, d.f, d.sigma, d.R, idxs)
    g2 = rg.backward2(layout, h, g, d.q, d.p, d.f, d.sigma, d.R, idxs)
    for b in range(r.B):
        qs, qe = int(r.q_off[b]), int(r.q_off[b + 1])
        ps, pe = int(r.p_off[b]), int(r.p_off[b + 1])
        c = r.cloud(b)
        ins = [t.double() for t in (c.q, c.p, c.f, c.sigma, c.R)]
        ib = idxs[qs:qe].cpu()[None]
        assert plain.rel_err(f_out[qs:qe].cpu()[None], plain.forward_given_idxs(*ins, ib)[0]) < 1e-5, b
        ref1 = plain.backward_closed(g[qs:qe].cpu().double()[None], *ins, ib)
        for n, x, y, rows in zip(("dq", "dp", "df", "dsigma", "dR"), g1, ref1,
                                 [(qs, qe)] + [(ps, pe)] * 4):
            assert plain.rel_err(x[rows[0]:rows[1]].cpu()[None], y) < 1e-5, (b, n)
        ref2 = plain.backward2_closed(h[qs:qe].cpu().double()[None], g[qs:qe].cpu().double()[None], *ins, ib)
        for n, x, y, rows in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), g2, ref2,
                                 [(qs, qe), (qs, qe)] + [(ps, pe)] * 4):
            assert plain.rel_err(x[rows[0]:rows[1]].cpu()[None], y) < 1e-5, (b, n)


def eikonal_step_ragged(params, layout, K, v):
    f_out, *_ = knn_aggregate_aniso_ragged(*params, layout, K)
    dq = torch.autograd.grad(f_out, params[0], v.expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    return (L,) + tuple(torch.autograd.grad(L, params))


def test_ragged_autograd_matches_dense_loop(cuda):
    r = small_ragged(3)
    layout = RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    params = [t.clone().requires_grad_(True) for t in (d.q, d.p, d.f, d.sigma, d.R)]
    out = eikonal_step_ragged(params, layout, r.K, d.v)
    for b in range(r.B):
        qs, qe = int(r.q_off[b]), int(r.q_off[b + 1])
        ps, pe = int(r.p_off[b]), int(r.p_off[b + 1])
        c = r.cloud(b).to(cuda)
        cp = [t.contiguous().clone().requires_grad_(True) for t in (c.q, c.p, c.f, c.sigma, c.R)]
        f_out, *_ = knn_aggregate_aniso(*cp, r.K)
        dq = torch.autograd.grad(f_out, cp[0], d.v.expand_as(f_out), create_graph=True)[0]
        L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
        grads = torch.autograd.grad(L, cp)
        for x, y, rows in zip(out[1:], grads, [(qs, qe)] + [(ps, pe)] * 4):
            assert plain.rel_err(x[rows[0]:rows[1]].cpu(), y[0].cpu()) < 1e-6, b


@pytest.mark.skip(reason="the one GPU run of this capture ended in a host segfault inside "
                         "torch.cuda.graph's capture_end (no GPU fault); cause not identified, so "
                         "graph capture of the step stays switched off")
@pytest.mark.parametrize("ragged", [False, True])
def test_graph_capture_replays_match_eager(cuda, ragged):
    """Capture the whole fwd + 1st + 2nd-order step; every replay (with eager work and
    synchronisation in between) must equal the eager step bit for bit."""
    if ragged:
        r = small_ragged(4)
        layout = RaggedLayout(r.q_off, r.p_off, cuda)
        d = r.to(cuda)
        params = [t.clone().requires_grad_(True) for t in (d.q, d.p, d.f, d.sigma, d.R)]

        def step():
            return eikonal_step_ragged(params, layout, r.K, d.v)
    else:
        b = workloads.clustered_batch(Q=20000, P=9000, K=16, F=32, seed=4).to(cuda)
        params = [t.clone().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]

        def step():
            f_out, *_ = knn_aggregate_aniso(*params, b.K)
            dq = torch.autograd.grad(f_out, params[0], b.v.expand_as(f_out), create_graph=True)[0]
            L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
            return (L,) + tuple(torch.autograd.grad(L, params))

    eager = [t.clone() for t in step()]
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        for _ in range(2):
            step()
    torch.cuda.current_stream().wait_stream(s)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static_out = step()
    for rep in range(4):
        graph.replay()
        torch.cuda.synchronize()
        for x, y in zip(static_out, eager):
            assert torch.equal(x, y), rep
        junk = torch.randn(1 << 20, device=cuda) * 2.0  # eager work between replays
        junk.sum().item()


def test_ragged_point_index(cuda):
    """A point index built over the packed clouds reproduces the per-call search."""
    r = small_ragged(6)
    layout = RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    offs = (layout.q_off, layout.p_off, layout.q_off_host, layout.p_off_host)
    index = _C.build_index(d.p, d.sigma, d.R, offsets=offs)
    ref = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K, tune={"search": "tiles"})
    got = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K, tune={"index": index})
    for x, y in zip(got, ref):
        assert torch.equal(x, y)
