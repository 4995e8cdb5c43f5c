"""GPU parity at the full benchmark configurations.

* clustered 1M x 64K (B=1, Q=2^20, P=2^16, K=16, F=64): one full launch of the forward,
  first and second order; ``idxs`` bit-exact on sampled query rows, ``f_out``, ``w_out``
  and every gradient within 1e-5 of the FP64 oracle on sampled query rows and on sampled
  point rows (random rows plus the heaviest CSR rows, whose oracle sums run over every
  query that lists them, accumulated in query chunks on the CPU).  Upstreams: the
  Eikonal read-out (broadcast) and a general N(0,1) upstream.
* ragged, 64 clouds (P_i log-uniform [1024, 32768], Q_i log-uniform [4096, 131072],
  K=16, F=32): the batched forward, first and second order equal the dense call run per
  cloud bit for bit; the three cheapest clouds are checked against the oracle in full.
* serving, 16 resident clouds of 65,536 points (F=64, K=16), requests of Q log-uniform
  in [256, 65536]: each result equals the same request served alone and the direct
  dense call bit for bit, and the oracle on sampled rows.

Bars: max|gpu - oracle| / max|oracle| per output, over the rows checked.
"""
from __future__ import annotations

import math

import pytest
import torch

from cuda_knn_aggregate import _C, workloads
from cuda_knn_aggregate import ragged as rg
from cuda_knn_aggregate.serving import Request, Server, serve_alone
from oracle import knn_ref, plain

pytestmark = pytest.mark.gpu

TOL32 = 1e-5


def report(name, got, ref, tol=TOL32):
    err = plain.rel_err(got.detach().cpu().double(), ref.detach().cpu().double())
    print(f"  {name}: rel err {err:.3e} (tol {tol:.0e})", flush=True)
    assert err <= tol, (name, err)


def _rows_of(t, rows):
    return t[:, rows].contiguous()


def _point_oracle(fn, point_rows, idxs_cpu, q, p, f, sigma, R, per_query, chunk=4096):
    """Point-side outputs (dp, df, dsigma, dR: the last four) of the closed form ``fn`` at
    ``point_rows``, summed over every query row that lists one of them (FP64, in query
    chunks).  ``per_query``: query-indexed upstreams (each [1, Q, ...]) passed before q."""
    want = torch.zeros(p.shape[1], dtype=torch.bool)
    want[point_rows] = True
    hit = want[idxs_cpu[0]].any(dim=1).nonzero().flatten()
    acc = None
    for s in range(0, hit.numel(), chunk):
        rows = hit[s:s + chunk]
        ups = [u[:, rows].double() for u in per_query]
        out = fn(*ups, q[:, rows].double(), p.double(), f.double(), sigma.double(), R.double(), idxs_cpu[:, rows])
        pts = [o[:, point_rows] for o in out[-4:]]
        acc = pts if acc is None else [a + b for a, b in zip(acc, pts)]
    return acc, hit.numel()


def test_clustered_full_launch_values_and_gradients(cuda):
    b = workloads.clustered(seed=0)
    B, Q, P, F, K = (b.shape[k] for k in ("B", "Q", "P", "F", "K"))
    d = b.to(cuda)
    cache = _C.new_cache(B, Q, P, torch.float32, cuda)
    f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(d.q, d.p, d.f, d.sigma, d.R, K, cache=cache)
    gen = torch.Generator().manual_seed(77)
    g_gen = torch.randn(B, Q, F, generator=gen)
    h = torch.randn(B, Q, 3, generator=gen)
    v_view = d.v.expand(B, Q, F)
    outs = {}
    for name, g in (("eikonal", v_view), ("general", g_gen.to(cuda))):
        outs[name] = (
            _C.knn_aggregate_aniso_backward(g, w_out, d.q, d.p, d.f, d.sigma, d.R, w_out, idxs, cache=cache),
            _C.knn_aggregate_aniso_backward_2nd(h.to(cuda), g, w_out, d.q, d.p, d.f, d.sigma, d.R, w_out, idxs,
                                                cache=cache),
        )
    torch.cuda.synchronize()
    idx = idxs.cpu()
    # sampled query rows, and point rows: random plus the heaviest CSR rows
    qrows = torch.randperm(Q, generator=gen)[:2048].sort().values
    deg = torch.bincount(idx.reshape(-1), minlength=P)
    heavy = deg.topk(8).indices
    prow = torch.unique(torch.cat([heavy, torch.randperm(P, generator=gen)[:40]]))
    print(f"  CSR degrees: mean {deg.float().mean():.1f}, max {int(deg.max())}, heaviest rows "
          f"{deg[heavy].tolist()}", flush=True)
    rd, ri = knn_ref.knn_ref(_rows_of(b.q, qrows), b.p, b.sigma, b.R, K)
    assert torch.equal(idx[:, qrows], ri), "idxs differ from the oracle on sampled rows"
    assert torch.equal(dists.cpu()[:, qrows], rd), "dists differ from the oracle on sampled rows"
    ins64 = [b.p.double(), b.f.double(), b.sigma.double(), b.R.double()]
    qs = _rows_of(b.q, qrows).double()
    iq = idx[:, qrows]
    rf, rw, _ = plain.forward_given_idxs(qs, *ins64, iq)
    report("f_out", f_out.cpu()[:, qrows], rf)
    report("w_out", w_out.cpu()[:, qrows], rw)
    for name in ("eikonal", "general"):
        g_cpu = b.v.expand(B, Q, F) if name == "eikonal" else g_gen
        (dq, dp, df, ds, dR), (dgf, dq2, dp2, df2, ds2, dR2) = outs[name]
        r1 = plain.backward_closed(g_cpu[:, qrows].double(), qs, *ins64, iq)
        report(f"{name} dq", dq.cpu()[:, qrows], r1[0])
        r2 = plain.backward2_closed(h[:, qrows].double(), g_cpu[:, qrows].double(), qs, *ins64, iq)
        report(f"{name} d grad_f_out", dgf.cpu()[:, qrows], r2[0])
        report(f"{name} dq (2nd)", dq2.cpu()[:, qrows], r2[1])
        pt1, nq1 = _point_oracle(plain.backward_closed, prow, idx, b.q, b.p, b.f, b.sigma, b.R, [g_cpu])
        print(f"  {name}: {prow.numel()} point rows, their sums over {nq1} query rows", flush=True)
        for n, x, y in zip(("dp", "df", "dsigma", "dR"), (dp, df, ds, dR), pt1):
            report(f"{name} {n}", x.cpu()[:, prow], y)
        pt2, _ = _point_oracle(plain.backward2_closed, prow, idx, b.q, b.p, b.f, b.sigma, b.R, [h, g_cpu])
        for n, x, y in zip(("dp", "df", "dsigma", "dR"), (dp2, df2, ds2, dR2), pt2):
            report(f"{name} {n} (2nd)", x.cpu()[:, prow], y)


def test_ragged_64_clouds_equal_per_cloud_and_oracle(cuda):
    r = workloads.ragged_batch(seed=0)
    layout = rg.RaggedLayout(r.q_off, r.p_off, cuda)
    d = r.to(cuda)
    F = d.f.shape[1]
    print(f"  {layout.B} clouds, {layout.Q} queries, {layout.P} points", flush=True)
    cache = layout.new_cache(torch.float32)
    f_out, w_out, dists, idxs = rg.forward(layout, d.q, d.p, d.f, d.sigma, d.R, r.K, cache=cache)
    g = d.v.expand(layout.Q, F)
    h = torch.randn(layout.Q, 3, generator=torch.Generator().manual_seed(5)).to(cuda)
    g1 = rg.backward(layout, g, d.q, d.p, d.f, d.sigma, d.R, idxs, cache=cache)
    g2 = rg.backward2(layout, h, g, d.q, d.p, d.f, d.sigma, d.R, idxs, cache=cache)
    torch.cuda.synchronize()
    qo, po = r.q_off.tolist(), r.p_off.tolist()
    mism = []
    for c in range(layout.B):
        qs, qe, ps, pe = qo[c], qo[c + 1], po[c], po[c + 1]
        cq = d.q[qs:qe][None].contiguous()
        cp, cf, csg, cR = (t[ps:pe][None].contiguous() for t in (d.p, d.f, d.sigma, d.R))
        o = _C.knn_aggregate_aniso_forward(cq, cp, cf, csg, cR, r.K)
        cv = d.v.expand(1, qe - qs, F)
        o1 = _C.knn_aggregate_aniso_backward(cv, None, cq, cp, cf, csg, cR, None, o[3], check=False)
        o2 = _C.knn_aggregate_aniso_backward_2nd(h[qs:qe][None].contiguous(), cv, None, cq, cp, cf, csg, cR, None,
                                                 o[3], check=False)
        got = [f_out, w_out, dists, idxs] + list(g1) + list(g2)
        ref = list(o) + list(o1) + list(o2)
        names = ["f_out", "w_out", "dists", "idxs", "dq", "dp", "df", "dsigma", "dR",
                 "dgf", "dq2", "dp2", "df2", "dsigma2", "dR2"]
        qside = {"f_out", "w_out", "dists", "idxs", "dq", "dgf", "dq2"}
        for n, x, y in zip(names, got, ref):
            a, e = (qs, qe) if n in qside else (ps, pe)
            if not torch.equal(x[a:e], y[0]):
                mism.append((c, n))
    assert not mism, mism[:10]
    # the three cheapest clouds against the oracle in full
    cost = [(qo[c + 1] - qo[c]) * (po[c + 1] - po[c]) for c in range(layout.B)]
    for c in sorted(range(layout.B), key=lambda i: cost[i])[:3]:
        qs, qe, ps, pe = qo[c], qo[c + 1], po[c], po[c + 1]
        cb = r.cloud(c)
        print(f"  oracle cloud {c}: Q={qe - qs} P={pe - ps}", flush=True)
        rd, ri = knn_ref.knn_ref(cb.q, cb.p, cb.sigma, cb.R, r.K)
        assert torch.equal(idxs.cpu()[qs:qe], ri[0]) and torch.equal(dists.cpu()[qs:qe], rd[0])
        ins = [t.double() for t in (cb.q, cb.p, cb.f, cb.sigma, cb.R)]
        rf, rw, _ = plain.forward_given_idxs(*ins, ri)
        report("f_out", f_out.cpu()[qs:qe], rf[0])
        gv = r.v.double().expand(1, qe - qs, F)
        r1 = plain.backward_closed(gv, *ins, ri)
        for n, x, y, (a, e) in zip(("dq", "dp", "df", "dsigma", "dR"), g1, r1,
                                   [(qs, qe)] + [(ps, pe)] * 4):
            report(n, x.cpu()[a:e], y[0])
        r2 = plain.backward2_closed(h.cpu()[qs:qe][None].double(), gv, *ins, ri)
        for n, x, y, (a, e) in zip(("dgf", "dq2", "dp2", "df2", "dsigma2", "dR2"), g2, r2,
                                   [(qs, qe)] * 2 + [(ps, pe)] * 4):
            report(n, x.cpu()[a:e], y[0])


def test_serving_16_resident_64k_clouds(cuda):
    clouds = []
    for c in range(16):
        b = workloads.clustered_batch(Q=1, P=65536, K=16, F=64, seed=200 + c)
        clouds.append((b.p[0], b.f[0], b.sigma[0], b.R[0]))
    v = torch.randn(64, generator=torch.Generator().manual_seed(9))
    server = Server(clouds, 16, v, device=cuda, streams=4, hbm_budget=8 << 30, max_batch_queries=1 << 17)
    gen = torch.Generator().manual_seed(3)
    lo, hi = math.log(256), math.log(65536)
    reqs, t = [], 0.0
    for i in range(24):
        c = int(torch.randint(16, (1,), generator=gen))
        Q = int(round(math.exp(lo + (hi - lo) * float(torch.rand(1, generator=gen)))))
        p = clouds[c][0]
        q = p[torch.randint(p.shape[0], (Q,), generator=gen)] + 0.01 * torch.randn(Q, 3, generator=gen)
        t += float(torch.empty(1).exponential_(2000.0, generator=gen))
        reqs.append(Request(c, q.pin_memory(), arrival=t, rid=i))
    server.replay(reqs)
    print(f"  {len(reqs)} requests in {server.batches} batches", flush=True)
    assert server.batches < len(reqs)
    for r in reqs[:6]:
        fa, na = serve_alone(server, r)
        assert torch.equal(r.f_out, fa) and torch.equal(r.normal, na), r.rid
        p, f, s, R = (x.to(cuda)[None].contiguous() for x in clouds[r.cloud])
        q = r.q.to(cuda)[None]
        f_out, w_out, _, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, 16)
        dq = _C.knn_aggregate_aniso_backward(server.v.expand(1, q.shape[1], 64), None, q, p, f, s, R, None, idxs,
                                             needs=(True, False, False, False, False), check=False)[0]
        assert torch.equal(r.f_out, f_out[0].cpu()) and torch.equal(r.normal, dq[0].cpu()), r.rid
        rows = torch.randperm(q.shape[1], generator=gen)[:256]
        cq = r.q[rows][None].contiguous()
        cp = [x[None] for x in clouds[r.cloud]]
        rd, ri = knn_ref.knn_ref(cq, cp[0], cp[2], cp[3], 16)
        assert torch.equal(idxs.cpu()[:, rows], ri), r.rid
        ins = [x.double() for x in (cq, cp[0], cp[1], cp[2], cp[3])]
        report(f"request {r.rid} (Q={r.q.shape[0]}) f_out", r.f_out[rows], plain.forward_given_idxs(*ins, ri)[0][0])
        rdq = plain.backward_closed(v.double().expand(1, 256, 64), *ins, ri)[0]
        report(f"request {r.rid} normal", r.normal[rows], rdq[0])
