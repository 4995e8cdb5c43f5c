"""GPU tests of the search tunables and the autotuner: every candidate launch (algorithm,
queries per wave, sub-tile records, super-tile size, occupancy target, brute-scan LDS
split) returns the default launch's neighbours bit for bit, and the tuner's choice comes
back from its cache file without timing again."""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import _C, autotune, workloads
from cuda_knn_aggregate import ragged as rg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("K", [16, 8, 12])
def test_every_candidate_same_neighbours_dense(cuda, K):
    b = workloads.clustered_batch(Q=20000, P=20000, K=K, F=4, seed=K).to(cuda)
    ref = _C.knn_aniso(b.q, b.p, b.sigma, b.R, K)
    for t in autotune.CANDIDATES:
        got = _C.knn_aniso(b.q, b.p, b.sigma, b.R, K, tune=t)
        assert torch.equal(got[1], ref[1]) and torch.equal(got[0], ref[0]), t
    # and through the forward (aggregation on the same neighbours)
    fo = _C.knn_aggregate_aniso_forward(b.q, b.p, b.f, b.sigma, b.R, K)
    for t in autotune.CANDIDATES[:11]:
        assert all(torch.equal(x, y) for x, y in
                   zip(_C.knn_aggregate_aniso_forward(b.q, b.p, b.f, b.sigma, b.R, K, tune=t), fo)), t


def test_every_candidate_same_neighbours_ragged(cuda):
    r = workloads.ragged_batch(B=12, p_range=(2000, 20000), q_range=(500, 20000), K=16, F=4, seed=3)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    ref = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K)
    for t in autotune.CANDIDATES:
        got = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K, tune=t)
        assert all(torch.equal(x, y) for x, y in zip(got, ref)), t


def test_tuner_cache_round_trip(cuda, tmp_path, monkeypatch):
    monkeypatch.setenv("AKNN_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_mem", {})
    b = workloads.clustered_batch(Q=30000, P=16384, K=16, F=4, seed=5).to(cuda)
    best = autotune.tuned(b.q, b.p, b.sigma, b.R, 16)
    assert best in [dict(t) for t in autotune.CANDIDATES]
    assert (tmp_path / "tune.json").exists()
    monkeypatch.setattr(autotune, "_mem", {})

    def no_timing(*a, **k):
        raise AssertionError("timed again instead of reading the cache")

    monkeypatch.setattr(autotune, "_time", no_timing)
    assert autotune.tuned(b.q, b.p, b.sigma, b.R, 16) == best
    ref = _C.knn_aniso(b.q, b.p, b.sigma, b.R, 16)
    got = _C.knn_aniso(b.q, b.p, b.sigma, b.R, 16, tune=best)
    assert torch.equal(got[1], ref[1])


def test_public_call_with_tuner(cuda, tmp_path, monkeypatch):
    """``knn_aggregate_aniso(..., tune="auto")`` runs the tuner's choice: same outputs and
    gradients as the default launch."""
    from cuda_knn_aggregate import knn_aggregate_aniso

    monkeypatch.setenv("AKNN_TUNE_CACHE", str(tmp_path / "tune.json"))
    monkeypatch.setattr(autotune, "_mem", {})
    b = workloads.clustered_batch(Q=20000, P=9000, K=16, F=8, seed=9).to(cuda)
    res = []
    for tune in (None, "auto"):
        ins = [t.clone().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
        out = knn_aggregate_aniso(*ins, 16, tune=tune)
        out[0].square().sum().backward()
        res.append([o.detach() for o in out] + [x.grad for x in ins])
    for x, y in zip(*res):
        assert torch.equal(x, y)
    with pytest.raises(ValueError):
        knn_aggregate_aniso(b.q, b.p, b.f, b.sigma, b.R, 16, tune="fast")


def _same(a, b):
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_persistent_search_same_bits(cuda):
    """Resident blocks over the work list give the one-block-per-tile launch's neighbours,
    with more work items than resident blocks: dense, split walk (8 ways), ragged."""
    b = workloads.clustered_batch(Q=600000, P=30000, K=16, F=4, seed=21).to(cuda)
    grid = _C.knn_aniso(b.q, b.p, b.sigma, b.R, 16, tune={"search": "tiles", "persistent": -1})
    pers = _C.knn_aniso(b.q, b.p, b.sigma, b.R, 16, tune={"search": "tiles", "persistent": 1})
    assert _same(grid, pers)
    b = workloads.clustered_batch(Q=65536, P=65536, K=16, F=4, seed=22).to(cuda)
    for ways in (8, 4):
        tg = {"search": "tiles", "persistent": -1, "point_split": ways}
        tp = dict(tg, persistent=1)
        assert _same(_C.knn_aniso(b.q, b.p, b.sigma, b.R, 16, tune=tg),
                     _C.knn_aniso(b.q, b.p, b.sigma, b.R, 16, tune=tp)), ways
    r = workloads.ragged_batch(seed=0)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    f1 = d.f[:, :4].contiguous()
    outs = [rg.forward(layout, d.q, d.p, f1, d.sigma, d.R, r.K, tune={"search": "tiles", "persistent": s})
            for s in (-1, 1, 0)]
    assert _same(outs[0], outs[1]) and _same(outs[0], outs[2])


def test_persistent_search_captured(cuda):
    """The persistent launch (its counters cleared by a kernel, not a memset) replays from a
    HIP graph with the eager bits, also after new query values."""
    r = workloads.ragged_batch(seed=1)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    q = d.q.clone()
    f1 = d.f[:, :4].contiguous()
    tune = {"search": "tiles", "persistent": 1}

    def fwd():
        return rg.forward(layout, q, d.p, f1, d.sigma, d.R, r.K, tune=tune)

    eager = [t.clone() for t in fwd()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(2):
            fwd()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        static = fwd()
    for _ in range(3):
        graph.replay()
        torch.cuda.synchronize()
        assert _same(static, eager)
        torch.randn(1 << 20, device=cuda).sum().item()
    with torch.no_grad():
        q.add_(0.001 * torch.randn(q.shape, generator=torch.Generator().manual_seed(3)).to(cuda))
    fresh = [t.clone() for t in fwd()]
    graph.replay()
    torch.cuda.synchronize()
    assert _same(static, fresh) and not torch.equal(fresh[3], eager[3])
