"""Continuous-batching search server over an HNSWIndex.

Requests (a query row with its own k and ef) arrive over time.  The server
owns `slots` launch slots, each with its own HIP stream and buffers for the
largest batch that fits its share of the device-memory budget.  Whenever a
slot is free and requests are waiting, every waiting request (up to the
slot's capacity) is admitted and searched in one launch of pann_hnsw_search
(per-query k and ef, so k = 10 and k = 100 requests share a launch) on that
slot's stream -- while the other slots' batches are still running, so work
from requests that arrive during a batch overlaps it on the device.  All
buffers are allocated when the server is created; admitting allocates
nothing.  Latency of a request = the time its batch's results are back in
host memory - its arrival.

Each slot's stream has its own launch state in libpann, so batches on
different slots do not share scratch, counters or status words.  The
reference has no server (qsearchAll batches, beamSearch.h:599-646; the Python
binding answers one call at a time); this is the serving harness its HNSW
search (HNSW.hpp:1568-1613) runs in here.
"""
from __future__ import annotations

import time
from dataclasses import dataclass

import numpy as np


@dataclass
class ServeResult:
    latency_s: np.ndarray      # per request
    batches: int
    mean_batch: float
    wall_s: float
    max_in_flight: int

    def summary(self):
        lat = self.latency_s * 1e3
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "qps": len(lat) / self.wall_s,
                "batches": self.batches, "mean_batch": self.mean_batch,
                "max_batches_in_flight": self.max_in_flight}


class _Slot:
    def __init__(self, B, dim, max_ef, qdt, dev):
        import torch
        self.stream = torch.cuda.Stream(device=dev)  # default priority
        self.q_dev = torch.empty((B, dim), dtype=qdt, device=dev)
        self.k_dev = torch.empty(B, dtype=torch.int32, device=dev)
        self.ef_dev = torch.empty(B, dtype=torch.int32, device=dev)
        self.oi = torch.empty((B, max_ef), dtype=torch.int32, device=dev)
        self.od = torch.empty((B, max_ef), dtype=torch.float32, device=dev)
        self.oc = torch.empty(B, dtype=torch.int32, device=dev)
        self.q_host = torch.empty((B, dim), dtype=qdt, pin_memory=True)
        self.ke_host = torch.empty((2, B), dtype=torch.int32, pin_memory=True)
        self.oi_host = torch.empty((B, max_ef), dtype=torch.int32, pin_memory=True)
        self.oc_host = torch.empty(B, dtype=torch.int32, pin_memory=True)
        self.done = torch.cuda.Event()
        self.busy = False
        self.lo = self.hi = 0


class BatchServer:
    def __init__(self, index, max_ef: int, budget_bytes: int, *, slots: int = 2,
                 store: str = "native", rerank: bool = False, summation: str = "tree",
                 device: int = 0, launch: dict | None = None):
        """launch: HNSWIndex.search_device launch options (e.g. the seen set
        and occupancy autotune.tune_hnsw chose); results do not depend on them."""
        import torch
        self.index = index
        self.max_ef = int(max_ef)
        self.store, self.rerank, self.summation = store, rerank, summation
        self.launch = dict(launch or {})
        dim = index.model.dim
        qdt = torch.float32 if index.base.dtype == np.float32 else torch.uint8
        qbytes = dim * (4 if qdt == torch.float32 else 1)
        # per admitted request: its query row (pinned + device), k, ef, and
        # max_ef ids + distances out (device + pinned), count
        per_req = 2 * qbytes + 2 * 4 + 2 * self.max_ef * 8 + 8
        self.max_batch = max(1, int(budget_bytes // (per_req * slots)))
        self.dev = torch.device("cuda", device)
        self.slots = [_Slot(self.max_batch, dim, self.max_ef, qdt, self.dev)
                      for _ in range(slots)]
        # one full-size batch per slot before serving: the library allocates a
        # stream's launch state and scratch at first use, which must not stall
        # a timed run
        B = self.max_batch
        z = np.zeros((B, dim), np.float32 if qdt == torch.float32 else np.uint8)
        one = np.ones(B, np.int32)
        for s in self.slots:
            self._launch(s, z, one, one, 0, B)
        for s in self.slots:
            self._retire(s)

    def _launch(self, s: _Slot, Q, ks, efs, lo, hi):
        import torch
        m = hi - lo
        s.q_host[:m].numpy()[:] = Q[lo:hi]
        s.ke_host[0, :m].numpy()[:] = ks[lo:hi]
        s.ke_host[1, :m].numpy()[:] = efs[lo:hi]
        with torch.cuda.stream(s.stream):
            s.q_dev[:m].copy_(s.q_host[:m], non_blocking=True)
            s.k_dev[:m].copy_(s.ke_host[0, :m], non_blocking=True)
            s.ef_dev[:m].copy_(s.ke_host[1, :m], non_blocking=True)
            self.index.search_device(s.q_dev[:m], 0, self.max_ef, s.oi[:m], s.od[:m], s.oc[:m],
                                     store=self.store, rerank=self.rerank,
                                     summation=self.summation, k_per_query=s.k_dev[:m],
                                     ef_per_query=s.ef_dev[:m], stream=s.stream, async_=True,
                                     **self.launch)
            s.oi_host[:m].copy_(s.oi[:m], non_blocking=True)
            s.oc_host[:m].copy_(s.oc[:m], non_blocking=True)
            s.done.record(s.stream)
        s.busy, s.lo, s.hi = True, lo, hi

    def _retire(self, s: _Slot):
        """Waits for the slot's batch; returns views of its results."""
        m = s.hi - s.lo
        s.done.synchronize()
        self.index.base.search_status(s.stream)  # raises on a kernel status
        s.busy = False
        return s.oi_host[:m].numpy().view(np.uint32), s.oc_host[:m].numpy()

    def search_all(self, Q, ks, efs):
        """Closed loop: every request at once, split over the slots."""
        n = len(Q)
        ids = np.zeros((n, self.max_ef), np.uint32)
        cnt = np.zeros(n, np.int32)
        i = 0
        while i < n or any(s.busy for s in self.slots):
            for s in self.slots:
                if s.busy:
                    ids[s.lo:s.hi], cnt[s.lo:s.hi] = self._retire(s)
                if i < n:
                    j = min(n, i + self.max_batch)
                    self._launch(s, Q, ks, efs, i, j)
                    i = j
        return ids, cnt

    def serve(self, Q, ks, efs, arrivals_s, keep_results: bool = False):
        """Open loop: request i arrives arrivals_s[i] seconds after the start
        (non-decreasing).  Returns a ServeResult (and the results)."""
        n = len(Q)
        lat = np.zeros(n)
        ids = np.zeros((n, self.max_ef), np.uint32) if keep_results else None
        cnt = np.zeros(n, np.int32) if keep_results else None
        i = nb = max_fly = 0
        t0 = time.perf_counter()
        while i < n or any(s.busy for s in self.slots):
            now = time.perf_counter() - t0
            progressed = False
            for s in self.slots:
                if s.busy and s.done.query():  # finished: results are in host memory
                    oi, oc = self._retire(s)
                    lat[s.lo:s.hi] = (time.perf_counter() - t0) - arrivals_s[s.lo:s.hi]
                    if keep_results:
                        ids[s.lo:s.hi], cnt[s.lo:s.hi] = oi, oc
                    progressed = True
            now = time.perf_counter() - t0
            j = int(np.searchsorted(arrivals_s, now, side="right"))
            for s in self.slots:
                if i < j and not s.busy:
                    e = min(j, i + self.max_batch)
                    self._launch(s, Q, ks, efs, i, e)
                    i = e
                    nb += 1
                    progressed = True
            max_fly = max(max_fly, sum(s.busy for s in self.slots))
            if not progressed and i < n and not any(s.busy for s in self.slots):
                wait = arrivals_s[i] - (time.perf_counter() - t0)
                if wait > 200e-6:
                    time.sleep(wait - 100e-6)
        wall = time.perf_counter() - t0
        res = ServeResult(lat, nb, n / max(nb, 1), wall, max_fly)
        return (res, ids, cnt) if keep_results else res


def poisson_arrivals(n: int, rate_qps: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_qps, n))
