"""GPU test of the allocation-free forward (``aknn_forward_noalloc``), called through the C
ABI: its outputs equal ``aknn_forward``'s bit for bit on consecutive calls that share the
library's per-device scratch (the second with a smaller batch), on a shape whose default
split walk does not fit the scratch, and it rejects a batch too large for the scratch
with a clear error before launching anything."""
from __future__ import annotations

import ctypes

import pytest
import torch

from cuda_knn_aggregate import _C, _lib, workloads

pytestmark = pytest.mark.gpu


def noalloc_forward(q, p, f, sigma, R, K, tune=None, cache=None):
    B, Q, P, F = q.shape[0], q.shape[1], p.shape[1], f.shape[2]
    dt = q.dtype
    dev = q.device
    f_out = torch.empty(B, Q, F, dtype=dt, device=dev)
    w_out = torch.empty(B, Q, K, dtype=dt, device=dev)
    dists = torch.empty(B, Q, K, dtype=dt, device=dev)
    idxs = torch.empty(B, Q, K, dtype=torch.int64, device=dev)
    dims = _C.make_dims(B, Q, P, F, K, dt)
    lib = _lib.load()
    tn = _C._tune(tune)
    cs = _C._cache_struct(cache)
    with torch.cuda.device(dev):
        _lib.check(lib.aknn_forward_noalloc(ctypes.byref(dims), _C._tp(tn), q.data_ptr(), p.data_ptr(), f.data_ptr(),
                                            sigma.data_ptr(), R.data_ptr(), f_out.data_ptr(), w_out.data_ptr(),
                                            dists.data_ptr(), idxs.data_ptr(), _C._cp(cs),
                                            torch.cuda.current_stream(dev).cuda_stream))
    return [f_out, w_out, dists, idxs]


def same(a, b):
    return all(torch.equal(x, y) for x, y in zip(a, b))


def test_noalloc_forward_matches_forward_on_consecutive_calls(cuda):
    cases = [
        ("uniform B=2 Q=12288 P=4096 K=16 F=64", workloads.bench_small(seed=1), None),
        ("uniform B=1 Q=700 P=300 K=8 F=32", workloads.uniform_batch(1, 700, 300, 8, 32, seed=2), None),
        ("clustered Q=20000 P=20000 K=16 F=16, brute", workloads.clustered_batch(Q=20000, P=20000, F=16, seed=3),
         {"search": "brute"}),
    ]
    for name, b, tune in cases:
        print(f"  case: {name}", flush=True)
        d = b.to(cuda)
        ref = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, b.K, tune=tune)
        got = noalloc_forward(d.q, d.p, d.f, d.sigma, d.R, b.K, tune=tune)
        torch.cuda.synchronize()
        assert same(got, ref), name


def test_noalloc_forward_with_cache_and_smaller_split(cuda):
    """64K queries on 64K points: the default split walk's lists need about 136 MiB, so the
    entry takes a smaller split; results and the filled cache equal aknn_forward's."""
    b = workloads.clustered_batch(Q=65536, P=65536, K=16, F=64, seed=4)
    d = b.to(cuda)
    c_ref = _C.new_cache(1, 65536, 65536, torch.float32, cuda)
    c_got = _C.new_cache(1, 65536, 65536, torch.float32, cuda)
    ref = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, 16, cache=c_ref)
    got = noalloc_forward(d.q, d.p, d.f, d.sigma, d.R, 16, cache=c_got)
    torch.cuda.synchronize()
    assert same(got, ref)
    assert same(c_got, c_ref)
    # and once more with a smaller batch on the same scratch
    s = workloads.readme(seed=5).to(cuda)
    assert same(noalloc_forward(s.q, s.p, s.f, s.sigma, s.R, s.K),
                _C.knn_aggregate_aniso_forward(s.q, s.p, s.f, s.sigma, s.R, s.K))


def test_noalloc_forward_rejects_batches_beyond_the_scratch(cuda):
    """About 4M queries need more than the scratch even unsplit: a clear error, raised on
    the host before any launch."""
    Q, P = 4 << 20, 4096
    q = torch.rand(1, Q, 3, device=cuda)
    b = workloads.uniform_batch(1, 1, P, 1, 1, seed=6).to(cuda)
    with pytest.raises(RuntimeError, match="bytes of scratch"):
        noalloc_forward(q, b.p, b.f, b.sigma, b.R, 1, tune={"search": "tiles"})
    torch.cuda.synchronize()
