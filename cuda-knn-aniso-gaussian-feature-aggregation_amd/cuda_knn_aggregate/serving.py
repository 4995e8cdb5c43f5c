"""Continuous-batching request loop over resident point clouds.

Clouds (points, features, Gaussians) stay resident in HBM, packed into one ragged point
set with one search index.  Each request asks for the aggregated features ``f_out`` of
its queries against one cloud and the gradient ``n = d(f_out . v)/dq`` (the SDF-normal
read-out, ``v`` fixed per server), i.e. the forward plus the query-side first-order
backward.

Scheduling (one host thread, non-blocking):

* arrived requests wait in FIFO order;
* a batch is formed whenever a stream slot is free: waiting requests join in arrival
  order, whatever their clouds, while the batch stays under ``max_batch_queries`` and
  the in-flight HBM estimate stays under ``hbm_budget`` (admission control);
* a batch is one ragged launch sequence (search + aggregation + dq, every cloud of the
  batch in the same launches, its queries grouped by cloud) on its own HIP stream, in
  buffers preallocated for that stream slot, so batches overlap on the device; the
  queries go in and the results come back through the slot's pinned staging buffers,
  moved by one copy kernel each way (no runtime copy calls on the scheduling thread,
  where they were seen to block for milliseconds), with a completion event;
* completed batches release their slot and budget at once, so new arrivals are admitted
  while other work is still on the GPU.

Each row's result does not depend on the batch it is served in (exact search, per-row
aggregation and gradient; a ragged launch gives each cloud's rows the bits of a dense
launch on that cloud alone), so a request gets the same bits alone and in any mix.
"""
from __future__ import annotations

import collections
import ctypes
import time
from dataclasses import dataclass
from typing import List, Optional

import torch

from . import _C, _lib


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
    row: int = 0                     # first row of its queries in the batch
    ready: Optional[torch.cuda.Event] = None  # device queries: recorded on their producer's stream

    @property
    def latency(self) -> float:
        return self.done - self.arrival


def _runs(rows):
    """Merge (start, length) row ranges into maximal contiguous runs."""
    out = []
    for a, m in sorted(rows):
        if out and out[-1][0] + out[-1][1] == a:
            out[-1] = (out[-1][0], out[-1][1] + m)
        else:
            out.append((a, m))
    return out


@dataclass
class _Batch:
    reqs: List[Request]
    event: torch.cuda.Event
    slot: int
    bytes: int
    nq: int = 0


class _Slot:
    """One stream slot: its HIP stream and the buffers of a batch of up to ``max_q``
    queries (device outputs and workspace, query offsets, pinned host staging with the
    device addresses the copy kernel uses)."""

    def __init__(self, device, dtype, F, K, max_q, n_clouds, ws_bytes):
        self.stream = torch.cuda.Stream(device=device)
        self.q = torch.empty(max_q, 3, dtype=dtype, device=device)
        self.f_out = torch.empty(max_q, F, dtype=dtype, device=device)
        self.w_out = torch.empty(max_q, K, dtype=dtype, device=device)
        self.dists = torch.empty(max_q, K, dtype=dtype, device=device)
        self.idxs = torch.empty(max_q, K, dtype=torch.int64, device=device)
        self.dq = torch.empty(max_q, 3, dtype=dtype, device=device)
        self.ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device=device)
        self.q_off = torch.zeros(n_clouds + 1, dtype=torch.int64, device=device)
        self.q_host = torch.empty(max_q, 3, dtype=dtype, pin_memory=True)
        self.q_off_host = torch.zeros(n_clouds + 1, dtype=torch.int64, pin_memory=True)
        self.f_host = torch.empty(max_q, F, dtype=dtype, pin_memory=True)
        self.n_host = torch.empty(max_q, 3, dtype=dtype, pin_memory=True)
        with torch.cuda.device(device):
            self.q_host_dev, self.q_off_host_dev, self.f_host_dev, self.n_host_dev = (
                _lib.host_device_pointer(t) for t in (self.q_host, self.q_off_host, self.f_host, self.n_host))
        self.event = torch.cuda.Event()


class Server:
    def __init__(self, clouds, K, v, *, device="cuda:0", streams=4, hbm_budget=2 << 30,
                 max_batch_queries=131072, tune=None, key="aniso", on_done=None):
        """clouds: list of (p [P,3], f [P,F], sigma [P,3], R [P,3,3]) tensors (moved to
        ``device``, packed and kept resident); v: [F] read-out for the normal.

        ``on_done(req)`` is called as each request completes, with ``req.f_out`` and
        ``req.normal`` as views into the slot's pinned staging buffers, valid until the
        callback returns (copy what must outlive it).  Without a callback the results
        are copied into each request.  Requests should hold their queries in pinned host
        memory (``tensor.pin_memory()``) so the uploads run asynchronously."""
        self.device = torch.device(device)
        self.K, self.key = int(K), key
        parts = [tuple(t.to(self.device).contiguous() for t in c) for c in clouds]
        if not parts:
            raise ValueError("no clouds")
        self.F = parts[0][1].shape[-1]
        self.dtype = parts[0][1].dtype
        for p, f, s, R in parts:
            if f.shape[-1] != self.F or f.dtype != self.dtype:
                raise ValueError("all clouds must share the feature width and dtype")
            if p.shape[0] < self.K:
                raise ValueError("every cloud needs at least K points")
        self.n_clouds = len(parts)
        self.p, self.f, self.sigma, self.R = (torch.cat([c[i] for c in parts]).contiguous() for i in range(4))
        sizes = torch.tensor([0] + [c[0].shape[0] for c in parts], dtype=torch.int64)
        self.p_off_host = torch.cumsum(sizes, 0).contiguous()
        self.p_off = self.p_off_host.to(self.device)
        self.P = int(self.p_off_host[-1])
        self.v = v.to(self.device, self.dtype)
        self.hbm_budget = int(hbm_budget)
        self.max_batch_queries = int(max_batch_queries)
        self._lib = _lib.load()
        # one search index for the packed clouds: the per-batch search skips the point
        # preparation (built with an empty query side; the index holds points only)
        q0 = torch.zeros(self.n_clouds + 1, dtype=torch.int64)
        self.index = _C.build_index(self.p, self.sigma, self.R, key=key,
                                    offsets=(q0.to(self.device), self.p_off, q0, self.p_off_host))
        self.tune = _C._tune(tune) or _lib.Tune()
        self.tune.index = self.index.data_ptr()
        # The read-out v is a broadcast upstream (one row per cloud): its point dot
        # products a[n] = v . f[n] are fixed for the resident clouds, so they are formed
        # once here and every batch's dq reads them (no feature rows gathered per batch).
        self.g_rows = self.v.expand(self.n_clouds, self.F).contiguous()
        self.dots = _C.point_dots(self.g_rows, self.f, self.p, self.sigma, self.R,
                                  offsets=(q0.to(self.device), self.p_off, q0, self.p_off_host))
        self.dq_cache = _C._cache_struct((None, None, None, self.dots))
        ws = self._ws_bytes(self.max_batch_queries)
        self.slots = [_Slot(self.device, self.dtype, self.F, self.K, self.max_batch_queries, self.n_clouds, ws)
                      for _ in range(streams)]
        self.streams = [sl.stream for sl in self.slots]
        self.free_slots = list(range(streams))
        self.inflight: List[_Batch] = []
        self.in_use = 0
        self.waiting: collections.deque = collections.deque()
        self.t0 = time.perf_counter()
        self.batches = 0
        self.on_done = on_done
        # The set-up above ran on the current stream; the slot streams do not wait for it
        # by themselves, so finish it before any batch can be launched.
        torch.cuda.current_stream(self.device).synchronize()

    def _dims(self, nq, q_off, q_off_host):
        return _C.make_dims(self.n_clouds, nq, self.P, self.F, self.K, self.dtype, self.key,
                            offsets=(q_off, self.p_off, q_off_host, self.p_off_host))

    def _ws_bytes(self, nq) -> int:
        """Forward workspace for nq queries, the largest over how they spread over the
        clouds (all in one cloud, or evenly)."""
        worst = 0
        for spread in ("one", "even"):
            counts = torch.zeros(self.n_clouds, dtype=torch.int64)
            if spread == "one":
                counts[0] = nq
            else:
                counts[:] = nq // self.n_clouds
                counts[0] += nq - int(counts.sum())
            qh = torch.cat([torch.zeros(1, dtype=torch.int64), torch.cumsum(counts, 0)])
            d = self._dims(nq, qh, qh)
            worst = max(worst, self._lib.aknn_forward_workspace_size(ctypes.byref(d)))
        return worst

    # ------------------------------------------------------------------ clock
    def now(self) -> float:
        return time.perf_counter() - self.t0

    # ------------------------------------------------------------------ requests
    def submit(self, req: Request) -> None:
        if not 0 <= req.cloud < self.n_clouds:
            raise ValueError(f"unknown cloud {req.cloud}")
        if req.q.dim() != 2 or req.q.shape[1] != 3 or req.q.shape[0] < 1:
            raise ValueError("request queries must have shape [Q, 3] with Q >= 1")
        if req.q.dtype != self.dtype:
            raise ValueError(f"request queries must be {self.dtype}")
        if req.q.shape[0] > self.max_batch_queries:
            raise ValueError(f"request of {req.q.shape[0]} queries exceeds max_batch_queries")
        if self.batch_bytes(req.q.shape[0]) > self.hbm_budget:
            raise ValueError("request alone exceeds the HBM budget")
        if req.q.is_cuda:  # device queries: the slot stream must see their producer's work
            req.ready = torch.cuda.Event()
            req.ready.record(torch.cuda.current_stream(req.q.device))
        elif not req.q.is_contiguous():
            req.q = req.q.contiguous()  # copied into the staging as one block of rows
        self.waiting.append(req)

    def batch_bytes(self, nq: int) -> int:
        """HBM estimate of one batch in flight: per-query inputs, outputs and search
        workspace (the clouds and their index are resident and counted apart)."""
        el = self.p.element_size()
        per_q = 3 * el + self.F * el * 2 + self.K * (2 * el + 8) + 3 * el + 64
        return nq * per_q + (1 << 20)

    # ------------------------------------------------------------------ engine
    def _launch(self, reqs: List[Request], slot: int, nbytes: int) -> None:
        sl = self.slots[slot]
        st = sl.stream
        t = self.now()
        # queries grouped by cloud (arrival order within a cloud): offsets and rows
        counts = [0] * self.n_clouds
        for r in reqs:
            counts[r.cloud] += r.q.shape[0]
        start = [0] * (self.n_clouds + 1)
        for c in range(self.n_clouds):
            start[c + 1] = start[c] + counts[c]
        nq = start[-1]
        offs = (ctypes.c_int64 * len(start))(*start)
        ctypes.memmove(sl.q_off_host.data_ptr(), offs, ctypes.sizeof(offs))
        fill = start[:-1]
        el = sl.q.element_size()
        host_rows = []
        for r in reqs:
            m = r.q.shape[0]
            r.started = t
            r.row = fill[r.cloud]
            fill[r.cloud] += m
            if r.q.is_cuda:  # device queries: a device copy after their producer's work
                with torch.cuda.stream(st):
                    st.wait_event(r.ready)
                    sl.q[r.row:r.row + m].copy_(r.q, non_blocking=True)
            else:  # host queries: into the slot's pinned staging by this thread (one memcpy)
                ctypes.memmove(sl.q_host.data_ptr() + r.row * 3 * el, r.q.data_ptr(), m * 3 * el)
                host_rows.append((r.row, m))
        # one copy kernel brings the staged rows and the offsets in (the slot's staging is
        # reused only after its previous batch completed)
        items = [(sl.q_off.data_ptr(), sl.q_off_host_dev, (self.n_clouds + 1) * 8)]
        for row, m in _runs(host_rows):
            items.append((sl.q.data_ptr() + row * 3 * el, sl.q_host_dev + row * 3 * el, m * 3 * el))
        _lib.copy(items, st.cuda_stream)
        # forward and the query-side dq through the C ABI on the slot's stream, into the
        # slot's buffers (inputs validated at construction and submission)
        lib = self._lib
        dims = self._dims(nq, sl.q_off, sl.q_off_host)
        ws_bytes = lib.aknn_forward_workspace_size(ctypes.byref(dims))
        if ws_bytes > sl.ws.numel():  # not expected: sized for the worst spread at start
            with torch.cuda.stream(st):
                sl.ws = torch.empty(ws_bytes, dtype=torch.uint8, device=self.device)
        hs = st.cuda_stream
        _lib.check(lib.aknn_forward(ctypes.byref(dims), ctypes.byref(self.tune), sl.q.data_ptr(),
                                    self.p.data_ptr(), self.f.data_ptr(), self.sigma.data_ptr(), self.R.data_ptr(),
                                    sl.f_out.data_ptr(), sl.w_out.data_ptr(), sl.dists.data_ptr(),
                                    sl.idxs.data_ptr(), sl.ws.data_ptr(), ws_bytes, None, hs))
        _lib.check(lib.aknn_backward(ctypes.byref(dims), _lib.NEED_DQ | _lib.GRAD_BROADCAST, self.g_rows.data_ptr(),
                                     sl.q.data_ptr(), self.p.data_ptr(), self.f.data_ptr(), self.sigma.data_ptr(),
                                     self.R.data_ptr(), sl.idxs.data_ptr(), None, None, ctypes.byref(self.dq_cache),
                                     None, 0, sl.dq.data_ptr(), None, None, None, None, hs))
        _lib.copy([(sl.f_host_dev, sl.f_out.data_ptr(), nq * self.F * el),
                   (sl.n_host_dev, sl.dq.data_ptr(), nq * 3 * el)], hs)
        sl.event.record(st)
        self.inflight.append(_Batch(reqs, sl.event, slot, nbytes, nq))
        self.in_use += nbytes
        self.batches += 1

    def _form_batches(self) -> None:
        while self.free_slots and self.waiting:
            take, nq = [], 0
            for r in self.waiting:
                n = r.q.shape[0]
                if take and (nq + n > self.max_batch_queries or
                             self.in_use + self.batch_bytes(nq + n) > self.hbm_budget):
                    break
                take.append(r)
                nq += n
            nbytes = self.batch_bytes(nq)
            if self.in_use + nbytes > self.hbm_budget:
                return  # wait for memory to free up
            for _ in take:
                self.waiting.popleft()
            self._launch(take, self.free_slots.pop(0), nbytes)

    def _reap(self) -> int:
        n = 0
        still = []
        for b in self.inflight:
            if b.event.query():
                t = self.now()
                sl = self.slots[b.slot]
                for r in b.reqs:
                    m = r.q.shape[0]
                    r.done = t
                    if self.on_done is not None:  # views, valid during the callback
                        r.f_out = sl.f_host[r.row:r.row + m]
                        r.normal = sl.n_host[r.row:r.row + m]
                        self.on_done(r)
                    else:
                        r.f_out = sl.f_host[r.row:r.row + m].clone()
                        r.normal = sl.n_host[r.row:r.row + m].clone()
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
        stream (first launches load code objects and size the workspaces)."""
        sizes = []
        n = 256
        while n < self.max_batch_queries:
            sizes.append(n)
            n *= 2
        sizes.append(self.max_batch_queries)
        keep = self.on_done
        self.on_done = lambda r: None
        try:
            for _ in range(repeats):
                for n in sizes:
                    for c in range(min(self.n_clouds, len(self.slots))):
                        ps, pe = int(self.p_off_host[c]), int(self.p_off_host[c + 1])
                        rows = torch.arange(n, device=self.device) % (pe - ps) + ps
                        self.submit(Request(c, self.p[rows].cpu().pin_memory()))
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
