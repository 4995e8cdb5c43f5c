"""GPU parity: the HIP library (through its C ABI, via the ``_C`` wrappers and the
autograd Functions) against the CPU oracle.

Bars (max|gpu - oracle| / max|oracle| per output):
* ``idxs`` and ``dists`` bit-exact against the FP32 oracle (same key expression,
  (key, index) order);
* ``f_out``, ``w_out`` and every first- and second-order gradient within 1e-5 of the
  FP64 oracle evaluated on the same neighbour indices;
* the FP64 build within 1e-10 of the FP64 oracle.
"""
from __future__ import annotations

import pytest
import torch

from cuda_knn_aggregate import KNNAnisotropicAggregate, _C, knn_aggregate_aniso, workloads
from oracle import knn_ref, plain

pytestmark = pytest.mark.gpu

TOL32 = 1e-5
TOL64 = 1e-10


def fwd(b, dev, key="aniso", tune=None):
    d = b.to(dev)
    return _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, b.K, key=key, tune=tune)


def oracle_knn(b, key="aniso"):
    return knn_ref.knn_ref(b.q, b.p, b.sigma, b.R, b.K, key=key)


def f64(b):
    return [t.double() for t in (b.q, b.p, b.f, b.sigma, b.R)]


def assert_rel(name, got, ref, tol):
    err = plain.rel_err(got.cpu(), ref)
    print(f"  {name}: rel err {err:.3e} (tol {tol:.0e})", flush=True)
    assert err <= tol, (name, err)


# ----------------------------------------------------------------------------- neighbours
def same_bits(a, b):
    """Equal values, NaN where the other has NaN."""
    a, b = a.cpu(), b.cpu()
    return torch.equal(a.isnan(), b.isnan()) and torch.equal(a[~a.isnan()], b[~b.isnan()])


def aniso_dists_at(b, idxs):
    """FP32 anisotropic keys of the given neighbours (the oracle's key expression)."""
    out = torch.empty(idxs.shape, dtype=b.q.dtype)
    for bi in range(b.q.shape[0]):
        D = plain.aniso_key(b.q[bi], b.p[bi], b.sigma[bi], b.R[bi])
        out[bi] = torch.gather(D, 1, idxs[bi])
    return out


@pytest.mark.parametrize("key", ["aniso", "euclidean"])
def test_knn_bitexact_readme(cuda, key):
    """Neighbours by the chosen key; dists are the anisotropic D either way."""
    b = workloads.readme()
    _, _, dists, idxs = fwd(b, cuda, key)
    rd, ri = oracle_knn(b, key)
    assert torch.equal(idxs.cpu(), ri)
    if key == "aniso":
        assert torch.equal(dists.cpu(), rd)
    assert torch.equal(dists.cpu(), aniso_dists_at(b, ri))


SEARCHES = ["brute", "tiles", "wave"]


@pytest.mark.parametrize("search", SEARCHES)
def test_knn_bitexact_bench_small(cuda, search):
    b = workloads.bench_small()
    _, _, dists, idxs = fwd(b, cuda, tune={"search": search})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri)
    assert torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("search", SEARCHES)
def test_knn_bitexact_clustered_full_launch(cuda, search):
    """The 1M-query / 64K-point launch; rows checked against the oracle on a sample."""
    b = workloads.clustered()
    _, _, dists, idxs = fwd(b, cuda, tune={"search": search})
    rows = torch.cat([torch.arange(0, 1536), torch.arange(1536, 1 << 20, 1021)])
    sub = workloads.Batch(b.q[:, rows], b.p, b.f, b.sigma, b.R, b.K)
    rd, ri = oracle_knn(sub)
    assert torch.equal(idxs.cpu()[:, rows], ri)
    assert torch.equal(dists.cpu()[:, rows], rd)


@pytest.mark.parametrize("search", ["tiles", "wave"])
@pytest.mark.parametrize("key", ["aniso", "euclidean"])
@pytest.mark.parametrize("K", [1, 7, 16, 40, 64])
def test_knn_tiles_k_and_key(cuda, K, key, search):
    b = workloads.clustered_batch(Q=3000, P=5000, K=K, F=4, seed=K)
    _, _, dists, idxs = fwd(b, cuda, key=key, tune={"search": search})
    rd, ri = oracle_knn(b, key)
    assert torch.equal(idxs.cpu(), ri)
    assert torch.equal(dists.cpu(), aniso_dists_at(b, ri))


@pytest.mark.parametrize("search", SEARCHES)
def test_knn_readme_literal_inputs(cuda, search):
    """The reference README's own inputs: R = torch.rand (not a rotation) and
    sigma = torch.rand (scales near 0 included)."""
    g = torch.Generator().manual_seed(21)
    B, Q, P, K, F = 2, 1000, 500, 12, 32
    b = workloads.Batch(torch.rand(B, Q, 3, generator=g), torch.rand(B, P, 3, generator=g),
                        torch.rand(B, P, F, generator=g), torch.rand(B, P, 3, generator=g),
                        torch.rand(B, P, 3, 3, generator=g), K)
    _, _, dists, idxs = fwd(b, cuda, tune={"search": search})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("search", ["tiles", "wave"])
def test_knn_tiles_degenerate_scales(cuda, search):
    """Zero and negative scales (tiles that must always be scanned), far outliers and
    exact duplicates."""
    b = workloads.clustered_batch(Q=2000, P=3000, K=16, F=4, seed=5)
    b.sigma[0, 100:140, 1] = 0.0
    b.sigma[0, 400:420, 2] = -0.01
    b.p[0, 900:905] += 50.0
    for t in (b.p, b.sigma, b.R):
        t[0, 2000:2400] = t[0, 1000:1400]
    _, _, dists, idxs = fwd(b, cuda, tune={"search": search})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("search,split", [("brute", 0), ("tiles", 0), ("tiles", 4), ("wave", 0)])
def test_knn_signed_zero_keys(cuda, search, split, dtype):
    """Keys of -0 (three points whose scales are all -0: every query's key to them is -0)
    tie with keys of +0 (queries placed on positive-scale points, some with lower
    indices): (key, index) order, and the sign bits of the keys exactly as scanned."""
    b = workloads.clustered_batch(Q=600, P=5000, K=16, F=4, seed=61)
    b = workloads.Batch(*(t.to(dtype) for t in (b.q, b.p, b.f, b.sigma, b.R)), b.K)
    b.sigma[0, 10:13] = -0.0
    b.q[0, 0:40] = b.p[0, 100:140]
    b.q[0, 40:60] = b.p[0, 0:20]
    tune = {"search": search}
    if split:
        tune["point_split"] = split
    _, _, dists, idxs = fwd(b, cuda, tune=tune)
    rd, ri = oracle_knn(b)
    bits = torch.int32 if dtype == torch.float32 else torch.int64
    neg0 = torch.tensor(-0.0, dtype=dtype).view(bits)
    assert bool((rd.view(bits) == neg0).sum(-1).ge(3).all())   # the case is exercised
    assert torch.equal(idxs.cpu(), ri)
    assert torch.equal(dists.cpu().view(bits), rd.view(bits))


@pytest.mark.parametrize("key", ["aniso", "euclidean"])
@pytest.mark.parametrize("ways", [1, 2, 4, 16, 64])
def test_knn_tiles_split_walk(cuda, ways, key):
    """Few, scattered queries: each query tile's walk split over `ways` blocks that
    share a pruning bound; same bits as the oracle whatever the split (ways above the
    library's cap for the shape are lowered to it)."""
    b = workloads.clustered_batch(Q=500, P=30000, K=16, F=4, seed=11)
    g = torch.Generator().manual_seed(12)
    b.q[0] = torch.rand(500, 3, generator=g) * 1.2 - 0.1      # spread over the whole cloud
    b.sigma[0, 100:140, 1] = 0.0
    b.sigma[0, 400:420, 2] = -0.01
    for t in (b.p, b.sigma, b.R):
        t[0, 20000:20400] = t[0, 1000:1400]                   # exact duplicates (ties)
    _, _, dists, idxs = fwd(b, cuda, key=key, tune={"search": "tiles", "point_split": ways})
    rd, ri = oracle_knn(b, key)
    assert torch.equal(idxs.cpu(), ri)
    assert torch.equal(dists.cpu(), aniso_dists_at(b, ri))


@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
@pytest.mark.parametrize("key", ["aniso", "euclidean"])
@pytest.mark.parametrize("search", ["auto", "wave"])
def test_knn_wave_sparse_queries(cuda, search, key, dtype):
    """Few queries spread over a cloud (the per-query walk's case, which AUTO picks
    here), with always-scanned tiles, exact duplicates and a NaN query; same bits as the
    oracle, in both precisions."""
    b = workloads.clustered_batch(Q=700, P=30000, K=16, F=4, seed=31)
    b = workloads.Batch(*(t.to(dtype) for t in (b.q, b.p, b.f, b.sigma, b.R)), b.K)
    g = torch.Generator().manual_seed(32)
    b.q[0] = torch.rand(700, 3, generator=g) * 1.2 - 0.1
    b.q[0, 5] = float("nan")
    b.sigma[0, 100:140, 1] = 0.0
    b.sigma[0, 400:420, 2] = -0.01
    for t in (b.p, b.sigma, b.R):
        t[0, 20000:20400] = t[0, 1000:1400]
    _, _, dists, idxs = fwd(b, cuda, key=key, tune={"search": search} if search != "auto" else None)
    rd, ri = oracle_knn(b, key)
    assert torch.equal(idxs.cpu(), ri)
    assert same_bits(dists, aniso_dists_at(b, ri))


@pytest.mark.parametrize("search", SEARCHES)
def test_knn_rows_without_k_comparable_keys(cuda, search):
    """A NaN query compares with no point: like the reference's scan it keeps points
    0..K-1, with NaN keys.  A cloud with fewer than K points whose keys are not NaN:
    those come first in (key, index) order and the remaining slots take the lowest
    unused indices.  Every index stays inside its cloud (the aggregation gathers by
    them)."""
    b = workloads.clustered_batch(Q=300, P=9000, K=16, F=4, seed=51)
    b.q[0, [0, 77, 299]] = float("nan")
    f_out, _, dists, idxs = fwd(b, cuda, tune={"search": search})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and same_bits(dists, rd)
    assert torch.equal(idxs.cpu()[0, 77], torch.arange(16)) and bool(f_out[0, 77].isnan().all())
    c = workloads.uniform_batch(1, 50, 40, 16, 4, seed=52)
    c.p[0, 5:35] = float("nan")                      # 10 points with comparable keys
    _, _, dists, idxs = fwd(c, cuda, tune={"search": search})
    idxs, dists = idxs.cpu()[0], dists.cpu()[0]
    finite = torch.tensor(list(range(5)) + list(range(35, 40)))
    for r in range(50):
        assert set(idxs[r, :10].tolist()) == set(finite.tolist())
        assert torch.equal(idxs[r, 10:], torch.tensor([5, 6, 7, 8, 9, 10]))
        assert bool(dists[r, 10:].isnan().all()) and not bool(dists[r, :10].isnan().any())


@pytest.mark.parametrize("shape", ["clustered", "uniform2"])
def test_knn_wave_large_cloud(cuda, shape):
    """Clouds of more than 64 super-tiles (the per-query walk orders them in chunks)."""
    if shape == "clustered":
        b = workloads.clustered_batch(Q=300, P=150000, K=16, F=4, seed=41)
    else:
        b = workloads.uniform_batch(2, 200, 140000, 16, 4, seed=42)
    _, _, dists, idxs = fwd(b, cuda, tune={"search": "wave"})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("tune", [{"sub_records": 8, "super_tiles": 16, "mu_classes": 2},
                                  {"sub_records": 4, "super_tiles": 8},
                                  {"point_split": 1},
                                  {"persistent": 1}])
def test_knn_tiles_walk_order_and_parts_equal_brute(cuda, tune):
    """131072 queries (2048 query tiles): the tiled search walks each XCD's tiles widest
    first and the widest of every 128 as four 16-query parts; every row equals the brute
    scan's (itself checked against the oracle above), in every launch mode."""
    b = workloads.clustered_batch(Q=131072, P=16384, K=16, F=1, seed=21, spread=0.05)
    d = b.to(cuda)
    rd, ri = _C.knn_aniso(d.q, d.p, d.sigma, d.R, b.K, tune={"search": "brute"})
    td, ti = _C.knn_aniso(d.q, d.p, d.sigma, d.R, b.K, tune=dict(tune, search="tiles"))
    assert torch.equal(ti, ri)
    assert torch.equal(td.view(torch.int32), rd.view(torch.int32))


@pytest.mark.parametrize("K", [1, 5, 12, 40])
@pytest.mark.parametrize("ways", [4, 64])
def test_knn_tiles_split_walk_k_values(cuda, ways, K):
    """Split walk with K below the kernel's list length (the shared bound and the seed
    are K-th keys, the lists keep more): same bits as the oracle."""
    b = workloads.clustered_batch(Q=400, P=30000, K=K, F=4, seed=60 + K)
    g = torch.Generator().manual_seed(61)
    b.q[0] = torch.rand(400, 3, generator=g) * 1.2 - 0.1
    _, _, dists, idxs = fwd(b, cuda, tune={"search": "tiles", "point_split": ways})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("qpt", [1, 2, 4])
@pytest.mark.parametrize("split", [1, 2, 4, 8])
def test_knn_launch_shapes(cuda, qpt, split):
    b = workloads.uniform_batch(2, 700, 900, 16, 8, seed=3)
    _, _, dists, idxs = fwd(b, cuda, tune={"queries_per_lane": qpt, "point_split": split})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


@pytest.mark.parametrize("K", [1, 2, 5, 12, 16, 17, 33, 64])
def test_knn_k_values(cuda, K):
    b = workloads.uniform_batch(1, 300, 200, K, 4, seed=K)
    _, _, dists, idxs = fwd(b, cuda)
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


def test_knn_k_equals_p(cuda):
    b = workloads.uniform_batch(2, 130, 16, 16, 4, seed=5)
    _, _, dists, idxs = fwd(b, cuda, tune={"point_split": 4})
    rd, ri = oracle_knn(b)
    assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd)


def test_knn_ties_duplicated_points(cuda):
    b = workloads.uniform_batch(1, 400, 300, 16, 4, seed=7)
    for t in (b.p, b.sigma, b.R):
        t[0, 150:300] = t[0, 0:150]
    rd, ri = oracle_knn(b)
    for tune in ({"point_split": 1}, {"point_split": 2}, {"point_split": 8}, {"search": "tiles"}):
        _, _, dists, idxs = fwd(b, cuda, tune=tune)
        assert torch.equal(idxs.cpu(), ri) and torch.equal(dists.cpu(), rd), tune


# ----------------------------------------------------------------------------- values + grads
def _check_all(b, dev, tol, dtype=torch.float32, key="aniso"):
    bd = b.to(dev, dtype)
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(bd.q, bd.p, bd.f, bd.sigma, bd.R, b.K, key=key)
    idx = idxs.cpu()
    ins = f64(b.to("cpu", dtype))
    rf, rw, rD = plain.forward_given_idxs(*ins, idx)
    assert_rel("f_out", f_out, rf, tol)
    assert_rel("w_out", w_out, rw, tol)
    g = torch.randn(b.q.shape[0], b.q.shape[1], b.f.shape[2], generator=torch.Generator().manual_seed(1))
    h = torch.randn(b.q.shape[0], b.q.shape[1], 3, generator=torch.Generator().manual_seed(2))
    g, h = g.to(dtype), h.to(dtype)
    gd, hd = g.to(dev), h.to(dev)
    got1 = _C.knn_aggregate_aniso_backward(gd, w_out, bd.q, bd.p, bd.f, bd.sigma, bd.R, w_out, idxs, key=key)
    ref1 = plain.backward_closed(g.double(), *ins, idx)
    for n, x, y in zip(("dq", "dp", "df", "dsigma", "dR"), got1, ref1):
        assert_rel(n, x, y, tol)
    got2 = _C.knn_aggregate_aniso_backward_2nd(hd, gd, w_out, bd.q, bd.p, bd.f, bd.sigma, bd.R, w_out, idxs, key=key)
    ref2 = plain.backward2_closed(h.double(), g.double(), *ins, idx)
    for n, x, y in zip(("dgrad_f_out", "dq2", "dp2", "df2", "dsigma2", "dR2"), got2, ref2):
        assert_rel(n, x, y, tol)
    return idx


def test_values_and_grads_readme(cuda):
    _check_all(workloads.readme(), cuda, TOL32)


def test_values_and_grads_sharp_sigma(cuda):
    """sigma in [5, 200]: weights spread over (0, 1] and gradients are large."""
    b = workloads.uniform_batch(2, 500, 400, 16, 64, seed=11, sigma_range=(5.0, 200.0))
    _check_all(b, cuda, TOL32)


def test_values_and_grads_bench_small(cuda):
    _check_all(workloads.bench_small(), cuda, TOL32)


@pytest.mark.parametrize("F", [1, 3, 16, 32, 48, 100, 128, 256])
def test_values_and_grads_feature_widths(cuda, F):
    b = workloads.uniform_batch(1, 200, 150, 8, F, seed=F, sigma_range=(1.0, 30.0))
    _check_all(b, cuda, TOL32)


def test_values_and_grads_euclidean_key(cuda):
    _check_all(workloads.readme(seed=4), cuda, TOL32, key="euclidean")


def test_fp64_path(cuda):
    b = workloads.uniform_batch(2, 300, 200, 12, 16, seed=9, sigma_range=(1.0, 50.0))
    b64 = b.to("cpu", torch.float64)
    idx = _check_all(b, cuda, TOL64, dtype=torch.float64)
    _, ri = knn_ref.knn_ref(b64.q, b64.p, b64.sigma, b64.R, b.K)
    assert torch.equal(idx, ri)


@pytest.mark.parametrize("key", ["aniso", "euclidean"])
def test_fp64_tiles(cuda, key):
    b = workloads.clustered_batch(Q=2000, P=3000, K=16, F=4, seed=6).to("cpu", torch.float64)
    _, _, dists, idxs = fwd(b, cuda, key=key, tune={"search": "tiles"})
    rd, ri = oracle_knn(b, key)
    assert torch.equal(idxs.cpu(), ri)


def test_deterministic_bits(cuda):
    b = workloads.bench_small(seed=2).to(cuda)
    outs = []
    for _ in range(2):
        f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(b.q, b.p, b.f, b.sigma, b.R, b.K)
        g1 = _C.knn_aggregate_aniso_backward(b.g, w_out, b.q, b.p, b.f, b.sigma, b.R, w_out, idxs)
        h = b.q * 0.5
        g2 = _C.knn_aggregate_aniso_backward_2nd(h, b.g, w_out, b.q, b.p, b.f, b.sigma, b.R, w_out, idxs)
        outs.append([f_out, w_out, dists, idxs] + g1 + g2)
    for x, y in zip(*outs):
        assert torch.equal(x, y)


# ----------------------------------------------------------------------------- autograd
def test_autograd_eikonal_step_matches_oracle(cuda):
    """fwd + 1st + 2nd order through the autograd Functions vs the plain FP64 autograd
    formulation on the same neighbours."""
    b = workloads.uniform_batch(2, 600, 300, 16, 32, seed=12, sigma_range=(2.0, 40.0))
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = knn_aggregate_aniso(*ins, b.K)
    v = b.v.to(cuda)
    dq = torch.autograd.grad(f_out, ins[0], v.expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    L.backward()

    def sel(q, p, s, R, K, key="aniso"):
        return None, idxs.cpu()

    Lr, gr = plain.eikonal_step(*f64(b), b.K, b.v.double(), select=sel)
    assert abs(L.item() - Lr.item()) <= TOL32 * abs(Lr.item())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, gr):
        assert_rel("grad " + n, x.grad, y, TOL32)


@pytest.mark.parametrize("F", [96, 127, 130, 256, 512, 1030])
def test_wide_feature_rows(cuda, F):
    """Feature rows held as 2 to 8 register chunks per 16-lane group (F = 96: two float4
    chunks; 127: eight scalar; 130: five float2 in the 8-chunk kernel; 512: the widest
    one call takes; 1030: three column ranges of 512, 512 and 6): aggregation, first
    order and the Eikonal step vs the FP64 oracle."""
    b = workloads.uniform_batch(1, 400, 300, 16, F, seed=70 + F, sigma_range=(2.0, 40.0))
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = knn_aggregate_aniso(*ins, b.K)
    ref_out = plain.forward_given_idxs(*f64(b), idxs.cpu())[0]
    assert_rel("f_out", f_out, ref_out, TOL32)
    (f_out * b.g.to(cuda)).sum().backward()
    ref = plain.backward_closed(b.g.double(), *f64(b), idxs.cpu())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref):
        assert_rel("grad " + n, x.grad, y, TOL32)
        x.grad = None
    f_out = knn_aggregate_aniso(*ins, b.K)[0]
    dq = torch.autograd.grad(f_out, ins[0], b.v.to(cuda).expand_as(f_out), create_graph=True)[0]
    L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
    L.backward()

    def sel(q, p, s, R, K, key="aniso"):
        return None, idxs.cpu()

    Lr, gr = plain.eikonal_step(*f64(b), b.K, b.v.double(), select=sel)
    assert abs(L.item() - Lr.item()) <= TOL32 * abs(Lr.item())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, gr):
        assert_rel("eikonal grad " + n, x.grad, y, TOL32)


def test_wide_feature_rows_fp64(cuda):
    """FP64 rows wider than one call (F = 300: column ranges of 256 and 44): forward and
    both backward orders vs the closed forms, to the FP64 tolerance."""
    b = workloads.uniform_batch(1, 300, 200, 8, 300, seed=91, sigma_range=(2.0, 40.0))
    x = [t.to(cuda, torch.float64) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(*x, b.K)
    ib = idxs.cpu()
    assert_rel("f_out", f_out, plain.forward_given_idxs(*f64(b), ib)[0], TOL64)
    g = b.g.double()
    h = torch.randn(1, 300, 3, dtype=torch.float64, generator=torch.Generator().manual_seed(92))
    got1 = _C.knn_aggregate_aniso_backward(g.to(cuda), None, *x, w_out, idxs)
    for n, u, v in zip(("dq", "dp", "df", "dsigma", "dR"), got1, plain.backward_closed(g, *f64(b), ib)):
        assert_rel(n, u, v, TOL64)
    got2 = _C.knn_aggregate_aniso_backward_2nd(h.to(cuda), g.to(cuda), None, *x, w_out, idxs)
    for n, u, v in zip(("dgf", "dq", "dp", "df", "dsigma", "dR"), got2, plain.backward2_closed(h, g, *f64(b), ib)):
        assert_rel(n, u, v, TOL64)


def test_autograd_first_order_matches_oracle(cuda):
    b = workloads.readme(seed=3)
    ins = [t.to(cuda).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    f_out, w_out, dists, idxs = knn_aggregate_aniso(*ins, b.K)
    (f_out * b.g.to(cuda)).sum().backward()
    ref = plain.backward_closed(b.g.double(), *f64(b), idxs.cpu())
    for n, x, y in zip(("q", "p", "f", "sigma", "R"), ins, ref):
        assert_rel("grad " + n, x.grad, y, TOL32)


def test_gradcheck_fp64_dq_path(cuda):
    """FP64 gradcheck + gradgradcheck of (q,p,f,sigma,R) -> d(f_out . v)/dq on the GPU,
    with K = P so the neighbour set cannot change under perturbation."""
    b = workloads.uniform_batch(1, 6, 5, 5, 3, seed=4, sigma_range=(1.0, 4.0))
    ins = [t.to(cuda, torch.float64).requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]
    v = torch.randn(3, dtype=torch.float64, device=cuda)

    def first(q, p, f, s, R):
        return knn_aggregate_aniso(q, p, f, s, R, 5)[0]

    def dq_path(q, p, f, s, R):
        out = knn_aggregate_aniso(q, p, f, s, R, 5)[0]
        return torch.autograd.grad(out, q, v.expand_as(out), create_graph=True)[0]

    assert torch.autograd.gradcheck(first, ins, eps=1e-6, atol=1e-8)
    assert torch.autograd.gradcheck(dq_path, ins, eps=1e-6, atol=1e-8)


# ----------------------------------------------------------------------------- errors
def test_errors(cuda):
    b = workloads.readme().to(cuda)
    with pytest.raises(RuntimeError, match="must not exceed"):
        knn_aggregate_aniso(b.q, b.p, b.f, b.sigma, b.R, 501)
    with pytest.raises(RuntimeError, match="contiguous"):
        knn_aggregate_aniso(b.q.transpose(0, 1), b.p, b.f, b.sigma, b.R, 4)
    with pytest.raises(RuntimeError, match="dtype"):
        knn_aggregate_aniso(b.q.double(), b.p, b.f, b.sigma, b.R, 4)
    with pytest.raises(RuntimeError, match="shape"):
        knn_aggregate_aniso(b.q, b.p, b.f, b.sigma[:, :10].contiguous(), b.R, 4)


@pytest.mark.parametrize("search", SEARCHES)
def test_cache_does_not_change_results(cuda, search):
    """The forward's cache (spatial row orders, packed point records) changes only the
    processing order of the gather kernels: every output is bit-identical."""
    b = workloads.clustered_batch(Q=5000, P=4000, K=16, F=32, seed=8, with_grad=True).to(cuda)
    args = (b.q, b.p, b.f, b.sigma, b.R)
    cache = _C.new_cache(1, 5000, 4000, torch.float32, cuda)
    r1 = _C.knn_aggregate_aniso_forward(*args, b.K, tune={"search": search})
    r2 = _C.knn_aggregate_aniso_forward(*args, b.K, tune={"search": search}, cache=cache)
    for x, y in zip(r1, r2):
        assert torch.equal(x, y)
    qo, po, rec = cache
    assert torch.equal(qo.sort().values.cpu(), torch.arange(5000, dtype=torch.int32))
    assert torch.equal(po.sort().values.cpu(), torch.arange(4000, dtype=torch.int32))
    w_out, idxs = r1[1], r1[3]
    g1 = _C.knn_aggregate_aniso_backward(b.g, w_out, *args, w_out, idxs)
    g2 = _C.knn_aggregate_aniso_backward(b.g, w_out, *args, w_out, idxs, cache=cache)
    for x, y in zip(g1, g2):
        assert torch.equal(x, y)
    h = torch.randn_like(b.q)
    s1 = _C.knn_aggregate_aniso_backward_2nd(h, b.g, w_out, *args, w_out, idxs)
    s2 = _C.knn_aggregate_aniso_backward_2nd(h, b.g, w_out, *args, w_out, idxs, cache=cache)
    for x, y in zip(s1, s2):
        assert torch.equal(x, y)


def _csr_expected(idxs, P):
    """Offsets and entries of the transpose, from torch's stable CPU sort of the global
    point rows (cloud b's rows start at b * P)."""
    B, Q, K = idxs.shape
    keys = (idxs + torch.arange(B).view(B, 1, 1) * P).reshape(-1)
    order = torch.sort(keys, stable=True).indices
    counts = torch.bincount(keys, minlength=B * P)
    off = torch.zeros(B * P + 1, dtype=torch.int64)
    off[1:] = torch.cumsum(counts, 0)
    return off, order


# Point keys of 9 to 16 bits with K divisible by 4 take the two packed passes (the
# 1M x 16 x 64K case fills all 32 bits of the packed word); 8-bit keys (one pass),
# 17-bit keys and K = 3 take the key and value pair sort.
@pytest.mark.parametrize("B,P,Q,K", [(1, 300, 700, 8), (1, 65536, 20000, 16), (1, 5000, 300000, 16),
                                     (3, 3000, 7000, 12), (1, 200, 1000, 4), (1, 70000, 5000, 16),
                                     (1, 70000, 40000, 16), (2, 5000, 3000, 3), (1, 65536, 1 << 20, 16)])
def test_csr_transpose(cuda, B, P, Q, K):
    """The point-side transpose of idxs: every entry once, each row's entries ascending
    (the fixed reduction order), offsets = exclusive scan of the per-point counts."""
    g = torch.Generator().manual_seed(P + Q)
    idxs = torch.randint(P, (B, Q, K), generator=g)
    idxs[0, : Q // 3, 0] = 7  # one heavy row
    off, ent = _C.build_csr(idxs.to(cuda), P)
    off_ref, ent_ref = _csr_expected(idxs, P)
    assert torch.equal(off.cpu().long(), off_ref)
    assert torch.equal(ent.cpu().long(), ent_ref)


@pytest.mark.parametrize("B,P,Q,K", [(1, 4000, 50000, 16), (2, 300, 2000, 8), (1, 200, 3000, 5)])
def test_csr_transpose_in_query_order(cuda, B, P, Q, K):
    """With a query order, entry q*K + k is numbered rank(q)*K + k (rank = the query's
    position in the order) and each row keeps the ascending q*K + k order."""
    g = torch.Generator().manual_seed(Q)
    idxs = torch.randint(P, (B, Q, K), generator=g)
    order = torch.randperm(B * Q, generator=g).to(torch.int32)
    off, ent, _ = _C.build_csr(idxs.to(cuda), P, order=order.to(cuda))
    off_ref, ent_ref = _csr_expected(idxs, P)
    rank = torch.empty(B * Q, dtype=torch.int64)
    rank[order.long()] = torch.arange(B * Q)
    numbered = rank[ent_ref // K] * K + ent_ref % K
    assert torch.equal(off.cpu().long(), off_ref)
    assert torch.equal(ent.cpu().long(), numbered)


@pytest.mark.parametrize("key", ["aniso", "euclidean"])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_point_index_gives_the_same_results(cuda, key, dtype):
    """A prebuilt point index (resident cloud) reused for several query sets: the
    same bits as the per-call preparation, small (split walk) and large query sets."""
    b = workloads.clustered_batch(Q=40000, P=20000, K=16, F=8, seed=41)
    b = workloads.Batch(*(t.to(dtype) for t in (b.q, b.p, b.f, b.sigma, b.R)), b.K).to(cuda)
    index = _C.build_index(b.p, b.sigma, b.R, key=key)
    for sl in (slice(0, 300), slice(0, 40000), slice(1000, 9000)):
        q = b.q[:, sl].contiguous()
        ref = _C.knn_aggregate_aniso_forward(q, b.p, b.f, b.sigma, b.R, 16, key=key, tune={"search": "tiles"})
        got = _C.knn_aggregate_aniso_forward(q, b.p, b.f, b.sigma, b.R, 16, key=key, tune={"index": index})
        for x, y in zip(got, ref):
            assert torch.equal(x, y)
