"""Continuous-batching request loop over resident point clouds.

Clouds (points, features, Gaussians) stay resident in HBM.  Each request asks for the
aggregated features ``f_out`` of its queries against one cloud and the gradient
``n = d(f_out . v)/dq`` (the SDF-normal read-out, ``v`` fixed per server), i.e. the
forward plus the query-side first-order backward.

Scheduling (one host thread, non-blocking):

* arrived requests wait in FIFO order;
* a batch is formed whenever a stream slot is free: the oldest waiting request picks
  the cloud, and later waiting requests for the same cloud join while the batch stays
  under ``max_batch_queries`` and the in-flight HBM estimate stays under
  ``hbm_budget`` (admission control);
* a batch is one dense launch sequence (search + aggregation + dq) on its own HIP
  stream, so batches for different clouds overlap on the device; results come back
  with an asynchronous device-to-host copy and a completion event;
* completed batches release their slot and budget at once, so new arrivals are admitted
  while other work is still on the GPU.

Each row's result does not depend on the batch it is served in (exact search, per-row
aggregation and gradient), so a request gets the same bits alone and in any mix.
"""
from __future__ import annotations

import collections
import time
from dataclasses import dataclass, field
from typing import List, Optional

import torch

from . import _C


@dataclass
class Request:
    cloud: int
    q: torch.Tensor                  # [Q, 3] host (pinned) or device
    arrival: float = 0.0             # seconds on the server clock
    started: float = -1.0            # launch of the batch that served it
    done: float = -1.0
    f_out: Optional[torch.Tensor] = None
    normal: Optional[torch.Tensor] = None
    rid: int = 0

    @property
    def latency(self) -> float:
        return self.done - self.arrival


@dataclass
class _Batch:
    reqs: List[Request]
    event: torch.cuda.Event
    slot: int
    bytes: int
    outs: list = field(default_factory=list)


class Server:
    def __init__(self, clouds, K, v, *, device="cuda:0", streams=4, hbm_budget=2 << 30,
                 max_batch_queries=131072, tune=None, key="aniso", on_done=None):
        """clouds: list of (p [P,3], f [P,F], sigma [P,3], R [P,3,3]) tensors (moved to
        ``device`` and kept resident); v: [F] read-out for the normal.  ``on_done(req)``
        is called as each request completes (a consumer that takes the results and
        drops them lets their pinned host buffers return to the allocator's cache)."""
        self.device = torch.device(device)
        self.K, self.key, self.tune = int(K), key, tune
        self.clouds = []
        for p, f, s, R in clouds:
            self.clouds.append(tuple(t.to(self.device).contiguous()[None] for t in (p, f, s, R)))
        self.F = self.clouds[0][1].shape[-1]
        self.v = v.to(self.device, self.clouds[0][1].dtype)
        # resident search indexes: the per-request search skips the point preparation
        self.indexes = [_C.build_index(p, s, R, key=key) for p, _, s, R in self.clouds]
        # the read-out row v for every query slot of the largest batch, built once
        self.g_all = self.v.expand(int(max_batch_queries), self.F).contiguous()
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(streams)]
        self.free_slots = list(range(streams))
        self.hbm_budget = int(hbm_budget)
        self.max_batch_queries = int(max_batch_queries)
        self.inflight: List[_Batch] = []
        self.in_use = 0
        self.waiting: collections.deque = collections.deque()
        self.t0 = time.perf_counter()
        self.batches = 0
        self.on_done = on_done

    # ------------------------------------------------------------------ clock
    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ------------------------------------------------------------------ requests
    def submit(self, req: Request) -> None:
        if not 0 <= req.cloud < len(self.clouds):
            raise ValueError(f"unknown cloud {req.cloud}")
        if req.q.dim() != 2 or req.q.shape[1] != 3 or req.q.shape[0] < 1:
            raise ValueError("request queries must have shape [Q, 3] with Q >= 1")
        if req.q.shape[0] > self.max_batch_queries:
            raise ValueError(f"request of {req.q.shape[0]} queries exceeds max_batch_queries")
        if self.batch_bytes(req.q.shape[0], req.cloud) > self.hbm_budget:
            raise ValueError("request alone exceeds the HBM budget")
        self.waiting.append(req)

    def batch_bytes(self, nq: int, cloud: int) -> int:
        """HBM estimate of one batch: per-query outputs, gradients and search workspace,
        plus the per-call point-side search structures of its cloud."""
        P = self.clouds[cloud][0].shape[1]
        el = self.clouds[cloud][0].element_size()
        per_q = 3 * el + self.F * el * 3 + self.K * (2 * el + 8) + 3 * el + 64
        return nq * per_q + P * (16 * el + 64) + (1 << 20)

    # ------------------------------------------------------------------ engine
    def _launch(self, reqs: List[Request], slot: int, nbytes: int) -> None:
        p, f, s, R = self.clouds[reqs[0].cloud]
        st = self.streams[slot]
        t = self.now()
        for r in reqs:
            r.started = t
        with torch.cuda.stream(st):
            qs = [r.q.to(self.device, non_blocking=True) for r in reqs]
            q = (torch.cat(qs, 0) if len(qs) > 1 else qs[0])[None].contiguous()
            tune = dict(self.tune or {})
            tune["index"] = self.indexes[reqs[0].cloud]
            f_out, w_out, dists, idxs = _C.knn_aggregate_aniso_forward(q, p, f, s, R, self.K, key=self.key,
                                                                       tune=tune)
            g = self.g_all[: q.shape[1]][None]
            (dq,) = _C.knn_aggregate_aniso_backward(g, w_out, q, p, f, s, R, w_out, idxs, key=self.key,
                                                   needs=(True, False, False, False, False))[:1]
            f_host = torch.empty(f_out.shape[1:], dtype=f_out.dtype, pin_memory=True)
            n_host = torch.empty(dq.shape[1:], dtype=dq.dtype, pin_memory=True)
            f_host.copy_(f_out[0], non_blocking=True)
            n_host.copy_(dq[0], non_blocking=True)
            ev = torch.cuda.Event()
            ev.record(st)
        b = _Batch(reqs, ev, slot, nbytes, [q, f_out, w_out, dists, idxs, g, dq, f_host, n_host])
        self.inflight.append(b)
        self.in_use += nbytes
        self.batches += 1

    def _form_batches(self) -> None:
        while self.free_slots and self.waiting:
            head = self.waiting[0]
            cloud = head.cloud
            take, nq = [], 0
            for r in list(self.waiting):
                if r.cloud != cloud:
                    continue
                n = r.q.shape[0]
                if take and (nq + n > self.max_batch_queries or
                             self.in_use + self.batch_bytes(nq + n, cloud) > self.hbm_budget):
                    break
                take.append(r)
                nq += n
            nbytes = self.batch_bytes(nq, cloud)
            if self.in_use + nbytes > self.hbm_budget:
                return  # wait for memory to free up
            for r in take:
                self.waiting.remove(r)
            self._launch(take, self.free_slots.pop(0), nbytes)

    def _reap(self) -> int:
        n = 0
        still = []
        for b in self.inflight:
            if b.event.query():
                t = self.now()
                f_host, n_host = b.outs[-2], b.outs[-1]
                off = 0
                for r in b.reqs:
                    m = r.q.shape[0]
                    r.f_out = f_host[off:off + m]
                    r.normal = n_host[off:off + m]
                    r.done = t
                    off += m
                    if self.on_done is not None:
                        self.on_done(r)
                self.free_slots.append(b.slot)
                self.in_use -= b.bytes
                n += 1
            else:
                still.append(b)
        self.inflight = still
        return n

    def poll(self) -> None:
        """One non-blocking scheduling round: reap finished batches, admit new ones."""
        self._reap()
        self._form_batches()

    def warm_up(self, repeats: int = 2) -> None:
        """Serve batches of every power-of-two size up to ``max_batch_queries`` on every
        stream, so that the device and pinned-host caching allocators hold blocks of
        each size class before timed serving starts."""
        sizes = []
        n = 256
        while n < self.max_batch_queries:
            sizes.append(n)
            n *= 2
        sizes.append(self.max_batch_queries)
        keep = self.on_done
        self.on_done = lambda r: (setattr(r, "f_out", None), setattr(r, "normal", None))
        try:
            for _ in range(repeats):
                for n in sizes:
                    for c in range(min(len(self.clouds), len(self.streams))):
                        p = self.clouds[c][0][0]
                        q = p[torch.arange(n, device=p.device) % p.shape[0]].cpu().pin_memory()
                        self.submit(Request(c, q))
                        self.poll()
                    self.drain()
        finally:
            self.on_done = keep

    def drain(self) -> None:
        while self.waiting or self.inflight:
            self.poll()

    # ------------------------------------------------------------------ replay
    def replay(self, requests: List[Request]) -> List[Request]:
        """Serve requests at their ``arrival`` times (seconds from now) and return them
        completed, with ``done`` set."""
        self.t0 = time.perf_counter()
        pending = collections.deque(sorted(requests, key=lambda r: r.arrival))
        while pending or self.waiting or self.inflight:
            t = self.now()
            while pending and pending[0].arrival <= t:
                self.submit(pending.popleft())
            self.poll()
        return requests


def latency_stats(reqs: List[Request]) -> dict:
    lat = sorted(r.latency for r in reqs)
    n = len(lat)

    def pct(x):
        return lat[min(n - 1, int(round(x / 100.0 * (n - 1))))]
    total_q = sum(r.q.shape[0] for r in reqs)
    span = max(r.done for r in reqs) - min(r.arrival for r in reqs)
    return {"requests": n, "p50_ms": pct(50) * 1e3, "p99_ms": pct(99) * 1e3, "max_ms": lat[-1] * 1e3,
            "queries_per_s": total_q / span if span > 0 else float("nan")}


def serve_alone(server: Server, req: Request):
    """One request through the same path with nothing else in flight (reference result)."""
    r = Request(req.cloud, req.q)
    server.submit(r)
    server.drain()
    return r.f_out, r.normal
