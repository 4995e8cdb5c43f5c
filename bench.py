"""Benchmark: fwd + 1st + 2nd-order (Eikonal) steps of knn_aggregate_aniso on one GPU.

One step = ``f_out = knn_aggregate_aniso(q, p, f, sigma, R, K)``; ``dq = d(f_out . v)/dq``
with ``create_graph=True``; ``L = sum_q (||dq||_2 - 1)^2``; gradients of ``L`` with respect
to q, p, f, sigma and R.  Default workload: one clustered cloud, Q = 2^20 queries,
P = 2^16 points, K = 16, F = 64, float32, synthetic seeded data, inputs resident in HBM.

Prints one JSON line: queries/s over the timed steps, the roofline of the dominant
kernel (per-launch hipEvent times recorded by the library's tracer over the timed
region, on the library's launch stream), the HBM roofline of the gather/scatter
kernels, the same step with a general (per-query N(0,1)) upstream, the reference's
published shape (B=2, Q=12288, P=4096, K=16, F=64: fwd, fwd+1st, fwd+1st+2nd), the
ragged batch against its per-cloud loop, the served latency at fixed Poisson rates,
and the plain-PyTorch CPU formulation timed on a bounded query sample on this host.
"""
from __future__ import annotations

import argparse
import hashlib
import json
import math
import os
import platform
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "cuda-knn-aniso-gaussian-feature-aggregation_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
KEY_FLOP = {"aniso": 26, "euclidean": 8}
METRIC = "fwd+bwd+2nd-order queries/s (Q=1M,P=64K,K=16,F=64) vs plain; served p50/p99 ms"
# compulsory bytes per query of aggregation + 1st + 2nd order with a materialised
# per-query upstream (int64 idxs, point arrays amortised over the queries), 1M x 64K
SURVEY_BYTES_PER_QUERY = 1706.75

WORKLOADS = {
    "clustered": "clustered B=1 Q=1048576 P=65536 K=16 F=64",
    "bench_small": "uniform B=2 Q=12288 P=4096 K=16 F=64",
    "readme": "uniform B=2 Q=1000 P=500 K=12 F=32",
    "ragged": "ragged 64 clouds P_i~logU[1024,32768] Q_i~logU[4096,131072] K=16 F=32",
    "serving": "serving 16 resident clouds P=65536 F=64 K=16; Poisson requests Q~logU[256,65536]; fwd + dq",
    "copy": "calibration: plain device copy of 1 GiB (float32), for HBM byte counters",
}
# the reference's published timings at B=2, Q=12288, P=4096, K=16, F=64 (README.md:47-55,
# one RTX 3080, labelled "per 100 iteration"; what "Backward" includes is not stated)
README_RTX3080_MS = {"fwd": 3.018, "bwd": 20.806, "bwd_2nd": 29.818}

# rocprofv3 PMC summary (FETCH_SIZE / WRITE_SIZE per launch, scaled by a plain-copy
# calibration) of the same workload, committed beside the kernel statistics
TRAFFIC_PROFILE = {"clustered": os.path.join(ROOT, "profiles", "hbm_traffic_clustered_1M_64K.md")}
LIB_SOURCES = os.path.join(PKG, "csrc")


def source_sha() -> str:
    """Hash of the HIP library's sources: a traffic profile applies to the build it was
    recorded on only."""
    h = hashlib.sha256()
    for name in sorted(os.listdir(LIB_SOURCES)):
        if name.endswith((".hip", ".h")):
            h.update(name.encode())
            with open(os.path.join(LIB_SOURCES, name), "rb") as fh:
                h.update(fh.read())
    for extra in (os.path.join(ROOT, "include", "aknn.h"),):
        with open(extra, "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(workload):
    """({kernel: (bytes read, bytes written) per launch}, source) from the committed
    profile; empty when the profile was recorded on other library sources."""
    path = TRAFFIC_PROFILE.get(workload)
    if not path or not os.path.exists(path):
        return {}, {"profile": None}
    out, inside, sha = {}, False, None
    for line in open(path):
        line = line.strip()
        if line.startswith("library sources:"):
            m = re.search(r"`([0-9a-f]+)`", line)
            sha = m.group(1) if m else line.split(":", 1)[1].strip()
        if line.startswith("```traffic"):
            inside = True
        elif line.startswith("```"):
            inside = False
        elif inside and line:
            name, rd, wr = line.rsplit(maxsplit=2)
            out[name] = (int(rd), int(wr))
    cur = source_sha()
    src = {"profile": os.path.relpath(path, ROOT), "profile_sources": sha, "current_sources": cur,
           "matches": sha == cur}
    return (out if sha == cur else {}), src


def stream_copy_gbs(dev, n=1 << 28, reps=5):
    """Device-to-device copy bandwidth on this GPU (1 GiB each way), GB/s."""
    import torch

    src = torch.empty(n, device=dev)
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(reps):
        dst.copy_(src)
    b.record()
    torch.cuda.synchronize()
    gbs = 2 * n * 4 * reps / (a.elapsed_time(b) * 1e-3) / 1e9
    del src, dst
    return gbs


def make_batch(name, seed=0):
    from cuda_knn_aggregate import workloads

    return {"clustered": workloads.clustered, "bench_small": workloads.bench_small,
            "readme": workloads.readme}[name](seed)


def algorithmic_bytes(kernel, B, Q, P, F, K, el=4, bcast=False):
    """Compulsory HBM bytes of one launch: inputs read once, outputs written once.  Point
    inputs p, sigma, R (15 values) and f (F values); int64 idxs.  ``bcast``: the upstream
    grad_f_out is one row per cloud (read once), not one per query."""
    rows_q, rows_p = B * Q, B * P
    pts_in = rows_p * (15 + F) * el
    pts_out = rows_p * (15 + F) * el
    g_in = B * F * el if bcast else rows_q * F * el
    if kernel == "aggregate":             # idxs, dists in; f_out, w_out out; f, q in
        return rows_q * (3 * el + 8 * K + el * K + F * el + el * K) + rows_p * F * el
    if kernel == "query_grad":            # g, q, idxs in; dq out; point inputs
        return g_in + rows_q * (3 * el + 8 * K + 3 * el) + pts_in
    if kernel == "query_grad2":           # h, g, q, idxs in; dq out
        return g_in + rows_q * (3 * el + 3 * el + 8 * K + 3 * el) + pts_in
    if kernel in ("dq_dots", "dq_dots2"):  # q, idxs (+h) in, dq out; point geometry + dots in
        return rows_q * ((6 if kernel == "dq_dots2" else 3) * el + 8 * K + 3 * el) + rows_p * 16 * el
    if kernel == "point_dots":            # f and the upstream rows in, one dot per point out
        return rows_p * F * el + B * F * el + rows_p * el
    if kernel == "point_grad":            # csr, q, g in; point grads out
        return rows_q * K * 4 + (rows_p + 1) * 4 + rows_q * 3 * el + g_in + pts_in + pts_out
    if kernel == "point_grad2":
        return rows_q * K * 4 + (rows_p + 1) * 4 + rows_q * 6 * el + g_in + pts_in + pts_out
    if kernel == "build_csr":             # idxs in, csr out
        return rows_q * K * (8 + 4) + (rows_p + 1) * 4
    if kernel in ("dq_reduce", "dq_reduce2"):  # per-entry shares in, dq out
        return rows_q * K * 3 * el + rows_q * 3 * el
    if kernel == "pack_points":
        return rows_p * (15 * el + 16 * el)
    if kernel == "knn_topk":              # q, records in; idxs, dists out
        return rows_q * (3 * el + 8 * K + el * K) + rows_p * 16 * el
    return 0


GATHER = ("aggregate", "query_grad", "query_grad2", "dq_dots", "dq_dots2", "point_dots", "point_grad",
          "point_grad2", "dq_reduce", "dq_reduce2")
SEARCH = ("knn_tiles_prep", "knn_tiles", "knn_wave", "knn_topk", "pack_points")


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def cpu_baseline(name, budget_s=15.0):
    """Plain-PyTorch CPU formulation (oracle.plain.eikonal_step, float32) on query samples
    of the same workload against all its points: the sample sized to about ``budget_s``
    of CPU work, a quarter of it, and twice it, to show the rate flat in the sample size."""
    import torch

    from oracle import plain

    torch.set_num_threads(max(1, min(os.cpu_count() or 1, int(os.environ.get("OMP_NUM_THREADS", "16")))))
    b = make_batch(name)
    Qtot = b.q.shape[1]

    def run(nq):
        rows = torch.arange(0, Qtot, max(1, Qtot // nq))[:nq]
        q = b.q[:, rows].contiguous()
        t0 = time.perf_counter()
        plain.eikonal_step(q, b.p, b.f, b.sigma, b.R, b.K, b.v)
        return time.perf_counter() - t0, q.shape[0] * q.shape[1]

    nq = min(256, Qtot)
    run(nq)                                        # warm-up
    t, n = run(nq)                                 # size probe (warm)
    nq = int(min(Qtot, max(nq, nq * budget_s / max(t, 1e-3) / 2)))
    t1, n1 = run(nq)
    t2, n2 = run(nq)
    t = min(t1, t2)
    tq, nq4 = run(max(1, nq // 4))                 # a quarter of the sample: the rate is linear in Q
    return {"value": n1 / t, "unit": "queries/s", "cores": torch.get_num_threads(), "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{n1} queries (every {max(1, Qtot // nq)}-th) of the {name} workload against all "
                      f"{b.p.shape[1]} points, plain-PyTorch autograd fwd+1st+2nd-order Eikonal step, "
                      f"float32, best of 2 warm runs ({t:.2f} s)",
            "quarter_sample": {"queries": nq4, "value": nq4 / tq},
            "rate_ratio_quarter_to_full": (nq4 / tq) / (n1 / t)}


def _timed(fn, steps, warm):
    import torch

    for _ in range(warm):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps


def eikonal_step_fn(params, K, up, tune=None):
    """One fwd + 1st (create_graph) + 2nd-order step; ``up(f_out)`` is the upstream of the
    first order (the read-out v broadcast, or a per-query tensor)."""
    import torch

    from cuda_knn_aggregate import knn_aggregate_aniso

    def step():
        f_out, w_out, dists, idxs = knn_aggregate_aniso(*params, K, tune=tune)
        dq = torch.autograd.grad(f_out, params[0], up(f_out), create_graph=True)[0]
        L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
        return torch.autograd.grad(L, params)

    return step


def captured_step_ms(step, steps, warm=2):
    """Capture ``step`` (library launches and elementwise torch ops only) in a HIP graph and
    time its replays; returns ms per replay."""
    import torch

    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        for _ in range(warm):
            step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        step()
    graph.replay()
    torch.cuda.synchronize()
    t = _timed(graph.replay, steps, 1)
    del graph
    return t * 1e3


def _loss_grad(dq):
    """h = d/d dq of sum (|dq| - 1)^2 by elementwise ops (capturable)."""
    import torch

    n = torch.sqrt(dq[..., 0:1] * dq[..., 0:1] + dq[..., 1:2] * dq[..., 1:2] + dq[..., 2:3] * dq[..., 2:3])
    return 2.0 * (n - 1.0) * dq / n


def explicit_step_fn(x, K, v, cache, tune=None):
    """The Eikonal step's gradients without autograd: forward, dq-only first order, the loss
    gradient h, and the second order of <h, dq> (the same launches the autograd step makes
    for these gradients), for capture."""
    from cuda_knn_aggregate import _C

    def step():
        q, p, f, s, R = x
        f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, K, cache=cache, tune=tune)
        up = v.expand_as(f_out)
        dq = _C.knn_aggregate_aniso_backward(up, None, q, p, f, s, R, None, idxs, cache=cache, check=False,
                                             needs=(True, False, False, False, False))[0]
        return _C.knn_aggregate_aniso_backward_2nd(_loss_grad(dq), up, None, q, p, f, s, R, None, idxs, cache=cache,
                                                   check=False, needs=(False, True, True, True, True, True))

    return step


def c1_times(steps=50, warm=5):
    """The reference's published shape (B=2, Q=12288, P=4096, K=16, F=64; README.md:47-55):
    forward, forward + first-order backward (upstream N(0,1) per query), and the Eikonal
    fwd + 1st + 2nd-order step, ms per iteration."""
    import torch

    from cuda_knn_aggregate import knn_aggregate_aniso

    dev = torch.device("cuda:0")
    b = make_batch("bench_small").to(dev)
    params = [t.clone().requires_grad_(True) for t in (b.q, b.p, b.f, b.sigma, b.R)]

    def fwd():
        with torch.no_grad():
            knn_aggregate_aniso(*params, b.K)

    def fwd_bwd():
        f_out = knn_aggregate_aniso(*params, b.K)[0]
        torch.autograd.grad(f_out, params, b.g)

    step = eikonal_step_fn(params, b.K, lambda f_out: b.v.expand_as(f_out))
    ms = {"fwd": _timed(fwd, steps, warm) * 1e3, "fwd_1st": _timed(fwd_bwd, steps, warm) * 1e3,
          "fwd_1st_2nd": _timed(step, steps, warm) * 1e3}
    return {"workload": WORKLOADS["bench_small"], "ms_per_iteration": ms,
            "queries_per_s": {k: 2 * 12288 / (v * 1e-3) for k, v in ms.items()},
            "iterations": steps, "upstream_1st": "N(0,1) per query", "upstream_eikonal": "v broadcast",
            "readme_rtx3080_ms": README_RTX3080_MS,
            "note": "README figures are labelled per 100 iterations; their Backward may include the forward"}


def ragged_summary(steps=10, warm=2, loop_steps=3):
    """Ragged batch: one batched fwd+1st+2nd-order step over all clouds (one launch per
    kernel) vs a per-cloud loop of the plain single-cloud call on the same GPU."""
    import torch

    from cuda_knn_aggregate import _lib, autotune, knn_aggregate_aniso, workloads
    from cuda_knn_aggregate.ragged import RaggedLayout, knn_aggregate_aniso_ragged

    dev = torch.device("cuda:0")
    r = workloads.ragged_batch(seed=0)
    layout = RaggedLayout(r.q_off, r.p_off, dev)
    d = r.to(dev)
    params = [t.clone().requires_grad_(True) for t in (d.q, d.p, d.f, d.sigma, d.R)]
    v = d.v
    tune = autotune.tuned(d.q, d.p, d.sigma, d.R, r.K, layout=layout)
    print(f"[bench] ragged: {r.B} clouds, {layout.Q} queries, {layout.P} points; search {tune}", flush=True)

    def batched():
        f_out, *_ = knn_aggregate_aniso_ragged(*params, layout, r.K, tune=tune)
        dq = torch.autograd.grad(f_out, params[0], v.expand_as(f_out), create_graph=True)[0]
        L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
        return torch.autograd.grad(L, params)

    clouds = []
    for b in range(r.B):
        c = r.cloud(b).to(dev)
        clouds.append([t.contiguous().clone().requires_grad_(True) for t in (c.q, c.p, c.f, c.sigma, c.R)])

    def per_cloud():
        for cp in clouds:
            f_out, *_ = knn_aggregate_aniso(*cp, r.K)
            dq = torch.autograd.grad(f_out, cp[0], v.expand_as(f_out), create_graph=True)[0]
            L = ((dq.norm(dim=-1) - 1.0) ** 2).sum()
            torch.autograd.grad(L, cp)

    with _lib.trace() as tr:
        tb = _timed(batched, steps, warm)
    tl = _timed(per_cloud, loop_steps, 1)
    kernels = {k: {"launches": n, "avg_ms": tot / n} for k, (n, tot) in tr.summary().items()}

    # the same gradients as one captured graph (forward, dq, loss gradient, second order)
    from cuda_knn_aggregate import ragged as rg

    x = [t.detach() for t in params]
    cache = layout.new_cache(x[0].dtype)

    def explicit():
        q, p, f, s, R = x
        f_out, _, _, idxs = rg.forward(layout, q, p, f, s, R, r.K, tune=tune, cache=cache)
        up = v.expand_as(f_out)
        dq = rg.backward(layout, up, q, p, f, s, R, idxs, cache=cache, needs=(True, False, False, False, False))[0]
        return rg.backward2(layout, _loss_grad(dq), up, q, p, f, s, R, idxs, cache=cache,
                            needs=(False, True, True, True, True, True))

    tg = captured_step_ms(explicit, steps)
    return {"workload": WORKLOADS["ragged"], "B": r.B, "Q_total": layout.Q, "P_total": layout.P, "K": r.K,
            "F": int(r.f.shape[1]), "value": layout.Q / tb, "unit": "queries/s", "ms_per_step": tb * 1e3,
            "steps": steps, "per_cloud_loop": {"value": layout.Q / tl, "ms_per_step": tl * 1e3},
            "batched_over_loop": tl / tb,
            "captured": {"ms_per_step": tg, "value": layout.Q / (tg * 1e-3), "over_loop": tl * 1e3 / tg,
                         "step": "one HIP graph: ragged forward, dq, loss gradient, second order"},
            "tune": tune, "kernels": kernels}


def _serving_setup():
    import torch

    from cuda_knn_aggregate import workloads
    from cuda_knn_aggregate.serving import Request, Server

    dev = torch.device("cuda:0")
    clouds = []
    for c in range(16):
        b = workloads.clustered_batch(Q=1, P=65536, K=16, F=64, seed=200 + c)
        clouds.append((b.p[0], b.f[0], b.sigma[0], b.R[0]))
    v = torch.randn(64, generator=torch.Generator().manual_seed(9))

    def consume(r):  # the client takes its results; only the timing is kept
        r.f_out = r.normal = None

    server = Server(clouds, 16, v, device=dev, streams=4, hbm_budget=8 << 30, max_batch_queries=1 << 17,
                    on_done=consume)

    def requests(n, rate, seed):
        g = torch.Generator().manual_seed(seed)
        out, t = [], 0.0
        lo, hi = math.log(256), math.log(65536)
        for i in range(n):
            c = int(torch.randint(16, (1,), generator=g))
            Q = int(round(math.exp(lo + (hi - lo) * float(torch.rand(1, generator=g)))))
            p = clouds[c][0]
            q = p[torch.randint(p.shape[0], (Q,), generator=g)] + 0.01 * torch.randn(Q, 3, generator=g)
            t += float(torch.empty(1).exponential_(rate, generator=g))
            out.append(Request(c, q.pin_memory(), arrival=t, rid=i))
        return out

    return server, requests


def _serve_at(server, requests, rate, n):
    import gc

    from cuda_knn_aggregate.serving import latency_stats

    reqs = requests(n, rate, int(rate))
    b0 = server.batches
    gc.collect()          # no collector pause inside the served interval
    gc.disable()
    try:
        server.replay(reqs)
    finally:
        gc.enable()
    st = latency_stats(reqs)
    st.update(rate_req_per_s=rate, batches=server.batches - b0,
              offered_queries_per_s=rate * sum(r.q.shape[0] for r in reqs) / len(reqs))
    return st


def served_summary(rates=(2000.0, 4000.0), n=1500):
    """Served latency at fixed Poisson rates (arrival to results in host memory)."""
    server, requests = _serving_setup()
    server.warm_up()
    server.replay(requests(100, 500.0, 1))
    rows = []
    for rate in rates:
        st = _serve_at(server, requests, rate, n)
        print(f"[bench] serving rate {rate}: {st}", flush=True)
        rows.append(st)
    return {"workload": WORKLOADS["serving"], "streams": 4, "rates": rows,
            "p99_under_5ms": [r["rate_req_per_s"] for r in rows if r["p99_ms"] < 5.0]}


def serving_bench(args):
    """Resident clouds, Poisson arrivals at several rates, continuous batching; p50/p99
    latency and served queries/s per rate."""
    server, requests = _serving_setup()
    server.warm_up()
    server.replay(requests(100, 500.0, 1))
    rows = []
    for rate in (250.0, 500.0, 1000.0, 2000.0, 3000.0, 4000.0, 5000.0, 6000.0, 8000.0):
        st = _serve_at(server, requests, rate, int(min(2000, max(200, rate * 2.0))))
        print(f"[bench] serving rate {rate}: {st}", flush=True)
        rows.append(st)
    best = max(rows, key=lambda r: r["queries_per_s"] if r["p99_ms"] < 5.0 else 0.0)
    out = {"metric": METRIC, "value": best["queries_per_s"], "unit": "queries/s", "n_gpus": 1,
           "steps": sum(r["requests"] for r in rows), "warmup": 100, "ms_per_step": best["p50_ms"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic", "config": {"workload": WORKLOADS["serving"], "streams": 4},
           "served_p50_ms": best["p50_ms"], "served_p99_ms": best["p99_ms"], "rates": rows}
    print(json.dumps(out), flush=True)


def ragged_bench(args):
    out = ragged_summary(args.steps, args.warmup, max(1, args.steps // 4))
    line = {"metric": METRIC, "value": out["value"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
            "warmup": args.warmup, "ms_per_step": out["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
            "config": {"workload": WORKLOADS["ragged"], "B": out["B"], "Q_total": out["Q_total"],
                       "P_total": out["P_total"], "K": out["K"], "F": out["F"],
                       "step": "fwd + 1st-order (create_graph) + 2nd-order Eikonal"}}
    line.update({k: out[k] for k in ("per_cloud_loop", "batched_over_loop", "tune", "kernels")})
    print(json.dumps(line), flush=True)


def copy_calibration(args):
    """A plain device-to-device copy of known size: reading a profiler's HBM byte
    counters on it gives the factor between counted and actual bytes."""
    import torch

    n = 1 << 28  # 1 GiB of float32 each way
    src = torch.rand(n, device="cuda:0")
    dst = torch.empty_like(src)
    for _ in range(args.warmup):
        dst.copy_(src)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        dst.copy_(src)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    print(json.dumps({"metric": "device copy GB/s", "value": 2 * n * 4 / dt / 1e9, "unit": "GB/s",
                      "bytes_read_per_launch": n * 4, "bytes_written_per_launch": n * 4,
                      "ms_per_step": dt * 1e3, "steps": args.steps, "warmup": args.warmup}), flush=True)


def _kernels(tr, steps):
    return {k: {"launches": n, "avg_ms": tot / n, "ms_per_step": tot / steps} for k, (n, tot) in tr.summary().items()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="clustered", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extras", action="store_true",
                    help="only the clustered step (no general-upstream step, reference shape, ragged, served)")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); the path shards by "
              "query across GPUs, which this build does not launch.", file=sys.stderr)
        sys.exit(2)

    if args.workload == "ragged":
        ragged_bench(args)
        return
    if args.workload == "serving":
        serving_bench(args)
        return
    if args.workload == "copy":
        copy_calibration(args)
        return

    import torch

    from cuda_knn_aggregate import _C, _lib, autotune

    dev = torch.device("cuda:0")
    b = make_batch(args.workload).to(dev)
    B, Q, P, F, K = (b.shape[k] for k in ("B", "Q", "P", "F", "K"))
    # the search launch the autotuner picks for this shape class (timed on first use,
    # outside the timed steps; the choice is kept in its cache file)
    tune = autotune.tuned(b.q, b.p, b.sigma, b.R, K)
    print(f"[bench] search tune {tune}", flush=True)
    params = [b.q, b.p, b.f, b.sigma, b.R]
    for t in params:
        t.requires_grad_(True)
    step = eikonal_step_fn(params, K, lambda f_out: b.v.expand_as(f_out), tune)

    print(f"[bench] workload {WORKLOADS[args.workload]}: warmup {args.warmup}", flush=True)
    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    print(f"[bench] timing {args.steps} steps", flush=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    value = B * Q / ((t1 - t0) / args.steps)
    # per-kernel times from a second pass of the same steps with the library's tracer on
    # (its hipEvents around every launch group are left out of the timed pass)
    with _lib.trace() as tr:
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(args.steps):
            step()
        torch.cuda.synchronize()
        t3 = time.perf_counter()
    traced_ms = (t3 - t2) * 1e3 / args.steps
    kernels = _kernels(tr, args.steps)
    dom = max(kernels, key=lambda k: kernels[k]["ms_per_step"])

    # (query, point) key evaluations of one search launch: all pairs for the brute scan,
    # counted by the tiled search's own device counters otherwise (untimed extra call)
    pairs = {"knn_topk": B * Q * P}
    if "knn_tiles" in kernels:
        st = torch.zeros(5, dtype=torch.int64, device=dev)
        _C.knn_aniso(b.q.detach(), b.p.detach(), b.sigma.detach(), b.R.detach(), K,
                     tune=dict(tune, stats=st))
        c = st.tolist()
        pairs["knn_tiles"] = c[3] * 64
        search_stats = {"waves": c[4], "super_tests_per_wave": c[0] / c[4], "tile_tests_per_wave": c[1] / c[4],
                        "tiles_scanned_per_wave": c[2] / c[4], "points_scanned_per_query": c[3] / c[4]}
    else:
        search_stats = None
    # degrees of the CSR point rows the point kernels walk (one wave per row): the skew
    # a heavy row puts on their tail
    _, idxs0 = _C.knn_aniso(b.q.detach(), b.p.detach(), b.sigma.detach(), b.R.detach(), K, tune=tune)
    off = _C.build_csr(idxs0, P)[0].long()
    deg = (off[1:] - off[:-1]).float()
    csr_degree = {"mean": float(deg.mean()), "max": int(deg.max()), "p99": float(torch.quantile(deg, 0.99)),
                  "p999": float(torch.quantile(deg, 0.999)), "empty_rows": int((deg == 0).sum()),
                  "entries_in_rows_over_4x_mean": float(deg[deg > 4 * deg.mean()].sum() / deg.sum())}
    del idxs0, off, deg

    traffic, traffic_src = measured_traffic(args.workload)

    def traffic_of(kname):
        t = traffic.get(kname)
        return None if t is None else t[0] + t[1]

    def roof(kname):
        avg_s = kernels[kname]["avg_ms"] * 1e-3
        if kname in pairs:
            flops = KEY_FLOP["aniso"] * pairs[kname]
            ach = flops / avg_s / 1e12
            return {"kernel": kname, "bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS,
                    "unit": "TFLOP/s", "frac": ach / FP32_PEAK_TFLOPS, "traffic": traffic_of(kname),
                    "traffic_source": traffic_src,
                    "per_launch": f"{KEY_FLOP['aniso']} FLOP x {pairs[kname]} evaluated (query, point) pairs",
                    "unfused_valu_peak": FP32_PEAK_TFLOPS / 4,
                    "frac_of_unfused_valu_peak": ach / (FP32_PEAK_TFLOPS / 4),
                    "brute_equivalent_tflops": KEY_FLOP["aniso"] * B * Q * P / avg_s / 1e12}
        byt = algorithmic_bytes(kname, B, Q, P, F, K, bcast=True)
        ach = byt / avg_s / 1e9
        return {"kernel": kname, "bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": traffic_of(kname), "traffic_source": traffic_src,
                "per_launch": f"{byt} algorithmic bytes"}

    def gather_roof(kern, steps, bcast):
        gather = [k for k in GATHER if k in kern]
        g_bytes = sum(algorithmic_bytes(k, B, Q, P, F, K, bcast=bcast) * kern[k]["launches"] / steps
                      for k in gather)
        g_s = sum(kern[k]["ms_per_step"] for k in gather) * 1e-3
        # the committed counters were recorded on the broadcast-upstream step only
        g_traffic = [traffic_of(k) if bcast else None for k in gather]
        # every launch of the step outside the neighbour search, against the compulsory
        # bytes per query of aggregation + 1st + 2nd order with a per-query upstream
        rest = [k for k in kern if k not in SEARCH]
        rest_s = sum(kern[k]["ms_per_step"] for k in rest) * 1e-3
        return {"kernels": gather, "bound": "hbm", "achieved": g_bytes / g_s / 1e9, "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": g_bytes / g_s / 1e9 / HBM_PEAK_GBS, "ms_per_step": g_s * 1e3,
                "bytes_per_query": g_bytes / (B * Q), "upstream": "broadcast row" if bcast else "per query",
                "traffic_per_step": None if (not gather or None in g_traffic) else sum(g_traffic),
                "all_non_search_kernels": {
                    "kernels": rest, "ms_per_step": rest_s * 1e3,
                    "survey_bytes_per_query": SURVEY_BYTES_PER_QUERY,
                    "frac": SURVEY_BYTES_PER_QUERY * B * Q / rest_s / 1e9 / HBM_PEAK_GBS}}

    roof_agg = gather_roof(kernels, args.steps, True)
    copy_gbs = stream_copy_gbs(dev)
    roof_agg.update(stream_copy_gbs=copy_gbs, frac_of_stream_copy=roof_agg["achieved"] / copy_gbs)

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": WORKLOADS[args.workload], "B": B, "Q": Q, "P": P, "K": K, "F": F,
                   "step": "fwd + 1st-order (create_graph, upstream v.expand_as(f_out)) + 2nd-order Eikonal "
                           "backward", "search_tune": tune},
        "roofline": roof(dom),
        "roofline_gather": roof_agg,
        "kernels": kernels,
        "kernels_pass_ms_per_step": traced_ms,
        "search_stats": search_stats,
        "csr_degree": csr_degree,
    }
    if not args.no_extras:
        # the step's gradients as one captured HIP graph (no autograd dispatch between launches)
        print("[bench] captured step", flush=True)
        xs = [t.detach() for t in params]
        cache = _C.new_cache(B, Q, P, xs[0].dtype, dev)
        tgraph = captured_step_ms(explicit_step_fn(xs, K, b.v, cache, tune), args.steps)
        out["captured_step"] = {"ms_per_step": tgraph, "value": B * Q / (tgraph * 1e-3), "unit": "queries/s",
                                "step": "one HIP graph: forward, dq (v broadcast), loss gradient, second order"}
        del cache
        # the same step with a general per-query upstream (materialised N(0,1) rows)
        print("[bench] general-upstream step", flush=True)
        g_up = torch.randn(B, Q, F, generator=torch.Generator().manual_seed(7)).to(dev)
        gstep = eikonal_step_fn(params, K, lambda f_out: g_up, tune)
        for _ in range(2):
            gstep()
        torch.cuda.synchronize()
        with _lib.trace() as tr2:
            tg = _timed(gstep, args.steps, 0)
        gk = _kernels(tr2, args.steps)
        out["general_upstream"] = {"value": B * Q / tg, "unit": "queries/s", "ms_per_step": tg * 1e3,
                                   "upstream": "N(0,1) per query, materialised [B, Q, F]",
                                   "roofline_gather": gather_roof(gk, args.steps, False), "kernels": gk}
        del g_up
        print("[bench] reference shape (B=2, Q=12288, P=4096)", flush=True)
        out["c1"] = c1_times()
        print("[bench] ragged batch", flush=True)
        out["ragged"] = ragged_summary()
        print("[bench] serving", flush=True)
        out["served"] = served_summary()
    if not args.no_cpu_baseline:
        print("[bench] cpu baseline", flush=True)
        out["cpu_baseline"] = cpu_baseline(args.workload, args.cpu_budget)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
