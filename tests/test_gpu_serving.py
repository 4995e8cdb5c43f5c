"""GPU tests of the continuous-batching server: every request's result equals the same
request served alone and the direct single call, whatever batch it shared."""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, workloads
from cuda_knn_aggregate.serving import Request, Server, latency_stats, serve_alone

pytestmark = pytest.mark.gpu


def make_server(cuda, n_clouds=3, P=5000, F=16, K=8, **kw):
    clouds = []
    for c in range(n_clouds):
        b = workloads.clustered_batch(Q=1, P=P, K=K, F=F, seed=100 + c)
        clouds.append((b.p[0], b.f[0], b.sigma[0], b.R[0]))
    v = torch.randn(F, generator=torch.Generator().manual_seed(5))
    return Server(clouds, K, v, device=cuda, **kw), clouds


def make_requests(clouds, n, seed=0, rate=2000.0, log_q=(5.5, 8.5)):
    g = torch.Generator().manual_seed(seed)
    reqs, t = [], 0.0
    for i in range(n):
        c = int(torch.randint(len(clouds), (1,), generator=g))
        Q = int(torch.exp(torch.empty(1).uniform_(*log_q, generator=g)).item())
        p = clouds[c][0]
        q = p[torch.randint(p.shape[0], (Q,), generator=g)] + 0.01 * torch.randn(Q, 3, generator=g)
        t += float(torch.empty(1).exponential_(rate, generator=g))
        reqs.append(Request(c, q.pin_memory(), arrival=t, rid=i))
    return reqs


def test_results_same_alone_and_in_the_mix(cuda):
    server, clouds = make_server(cuda, streams=3, max_batch_queries=6000)
    # arrivals within a few microseconds of each other: requests must share batches
    # (at a realistic rate the batches may finish before the next request arrives)
    reqs = make_requests(clouds, 40, rate=1e6)
    server.replay(reqs)
    assert all(r.done >= r.arrival for r in reqs)
    assert server.batches < len(reqs)          # some requests shared a batch
    for r in reqs[:12]:
        fa, na = serve_alone(server, r)
        assert torch.equal(r.f_out, fa) and torch.equal(r.normal, na)
        p, f, s, R = (t.to(cuda)[None].contiguous() for t in clouds[r.cloud])
        q = r.q.to(cuda)[None]
        f_out, w_out, _, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, server.K)
        # the read-out as a broadcast upstream, as the server takes it (dq summed from the
        # point dot products v . f[n]: the same bits in a dense call on the cloud alone)
        g = server.v.expand(1, q.shape[1], server.F)
        dq = _C.knn_aggregate_aniso_backward(g, w_out, q, p, f, s, R, w_out, idxs,
                                             needs=(True, False, False, False, False))[0]
        assert torch.equal(r.f_out, f_out[0].cpu()) and torch.equal(r.normal, dq[0].cpu())
    st = latency_stats(reqs)
    assert st["requests"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0


def test_admission_limits(cuda):
    server, clouds = make_server(cuda, hbm_budget=8 << 20, max_batch_queries=4096)
    with pytest.raises(ValueError):
        server.submit(Request(0, torch.rand(5000, 3)))      # above max_batch_queries
    with pytest.raises(ValueError):
        server.submit(Request(7, torch.rand(10, 3)))        # unknown cloud
    reqs = make_requests(clouds, 20, seed=3, rate=1e6, log_q=(5.0, 8.0))  # all arrive at once
    server.replay(reqs)
    assert all(r.done > 0 for r in reqs)
    assert server.in_use == 0 and not server.inflight
