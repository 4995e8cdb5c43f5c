"""HIP graph capture of the fwd + 1st + 2nd-order Eikonal step through the C ABI.

The step is issued from the calling thread: the library's forward, its dq-only first
order (upstream v broadcast), the loss gradient h = 2 (|dq| - 1) dq / |dq| by elementwise
torch ops, and the second order of <h, dq> -- every library launch is a kernel launch on
the capturing stream (no memset, copy, event or synchronisation).  Every replay must
equal the eager step bit for bit, also after the captured inputs are overwritten with
new values between replays, with eager work and synchronisations in between.
"""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, workloads
from cuda_knn_aggregate import ragged as rg

pytestmark = pytest.mark.gpu

DQ = (True, False, False, False, False)
ALL2 = (False, True, True, True, True, True)


def _loss_grad(dq):
    """h = d/d dq of sum (|dq| - 1)^2, elementwise only."""
    n = torch.sqrt(dq[..., 0:1] * dq[..., 0:1] + dq[..., 1:2] * dq[..., 1:2] + dq[..., 2:3] * dq[..., 2:3])
    return 2.0 * (n - 1.0) * dq / n


def dense_step(x, K, v, cache):
    q, p, f, s, R = x
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, K, cache=cache)
    up = v.expand_as(f_out)
    dq = _C.knn_aggregate_aniso_backward(up, None, q, p, f, s, R, None, idxs, needs=DQ, cache=cache, check=False)[0]
    h = _loss_grad(dq)
    g2 = _C.knn_aggregate_aniso_backward_2nd(h, up, None, q, p, f, s, R, None, idxs, needs=ALL2, cache=cache,
                                             check=False)
    return (f_out, dq) + tuple(g2[1:])


def ragged_step(x, layout, K, v, cache):
    q, p, f, s, R = x
    f_out, w_out, dists, idxs = rg.forward(layout, q, p, f, s, R, K, cache=cache)
    up = v.expand_as(f_out)
    dq = rg.backward(layout, up, q, p, f, s, R, idxs, needs=DQ, cache=cache)[0]
    h = _loss_grad(dq)
    g2 = rg.backward2(layout, h, up, q, p, f, s, R, idxs, needs=ALL2, cache=cache)
    return (f_out, dq) + tuple(g2[1:])


def _check_capture(cuda, x, step, perturb):
    eager = [t.clone() for t in step()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):  # warm-up: code objects and the allocator pool
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static_out = step()
    torch.cuda.synchronize()
    print("  captured", flush=True)
    for rep in range(3):
        graph.replay()
        torch.cuda.synchronize()
        for i, (a, b) in enumerate(zip(static_out, eager)):
            assert torch.equal(a, b), (rep, i)
        junk = torch.randn(1 << 20, device=cuda) * 2.0  # eager work between replays
        junk.sum().item()
    # new input values in the captured buffers: the replay recomputes everything
    with torch.no_grad():
        x[0].add_(perturb)
    fresh = [t.clone() for t in step()]
    graph.replay()
    torch.cuda.synchronize()
    for i, (a, b) in enumerate(zip(static_out, fresh)):
        assert torch.equal(a, b), ("perturbed", i)
    assert not torch.equal(fresh[0], eager[0])


def test_capture_dense_step(cuda):
    b = workloads.clustered_batch(Q=20000, P=9000, K=16, F=32, seed=4).to(cuda)
    x = [b.q.clone(), b.p, b.f, b.sigma, b.R]
    cache = _C.new_cache(1, 20000, 9000, torch.float32, cuda)
    perturb = 0.003 * torch.randn(x[0].shape, generator=torch.Generator().manual_seed(1)).to(cuda)
    _check_capture(cuda, x, lambda: dense_step(x, b.K, b.v, cache), perturb)


def test_capture_ragged_step(cuda):
    r = workloads.ragged_small(seed=4)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    x = [d.q.clone(), d.p, d.f, d.sigma, d.R]
    cache = layout.new_cache(torch.float32)
    perturb = 0.01 * torch.randn(x[0].shape, generator=torch.Generator().manual_seed(2)).to(cuda)
    _check_capture(cuda, x, lambda: ragged_step(x, layout, r.K, d.v, cache), perturb)
