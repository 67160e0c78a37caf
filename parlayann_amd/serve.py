"""Continuous-batching search server over an HNSWIndex.

Requests (a query row with its own k and ef) arrive over time.  Whenever the
device is free, every request that has arrived is admitted, up to the batch
that fits the device-memory budget, and searched in one launch of
pann_hnsw_search (per-query k and ef, so k = 10 and k = 100 requests share a
launch).  Buffers for the largest admissible batch are allocated once, when
the server is created; a round allocates nothing.  Latency of a request =
completion of its batch (results back in host memory) - its arrival.

The reference has no server (queries are batched by qsearchAll,
beamSearch.h:599-646, or answered one by one by the Python binding); this
is the serving harness its HNSW search (HNSW.hpp:1568-1613) runs in here.
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

    def summary(self):
        lat = self.latency_s * 1e3
        return {"p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
                "mean_ms": float(lat.mean()), "qps": len(lat) / self.wall_s,
                "batches": self.batches, "mean_batch": self.mean_batch}


class BatchServer:
    def __init__(self, index, max_ef: int, budget_bytes: int, *, store: str = "native",
                 rerank: bool = False, summation: str = "tree", device: int = 0):
        import torch
        self.index = index
        self.max_ef = int(max_ef)
        self.store, self.rerank, self.summation = store, rerank, summation
        dim = index.model.dim
        qdt = torch.float32 if index.base.dtype == np.float32 else torch.uint8
        qbytes = dim * (4 if qdt == torch.float32 else 1)
        # per admitted request: its query row (pinned + device), k, ef, and
        # max_ef ids + distances out (device + pinned), count
        per_req = 2 * qbytes + 2 * 4 + 2 * self.max_ef * 8 + 8
        self.max_batch = max(1, int(budget_bytes // per_req))
        self.dev = torch.device("cuda", device)
        B = self.max_batch
        self.q_dev = torch.empty((B, dim), dtype=qdt, device=self.dev)
        self.k_dev = torch.empty(B, dtype=torch.int32, device=self.dev)
        self.ef_dev = torch.empty(B, dtype=torch.int32, device=self.dev)
        self.oi = torch.empty((B, self.max_ef), dtype=torch.int32, device=self.dev)
        self.od = torch.empty((B, self.max_ef), dtype=torch.float32, device=self.dev)
        self.oc = torch.empty(B, dtype=torch.int32, device=self.dev)
        self.q_host = torch.empty((B, dim), dtype=qdt, pin_memory=True)
        self.ke_host = torch.empty((2, B), dtype=torch.int32, pin_memory=True)
        self.oi_host = torch.empty((B, self.max_ef), dtype=torch.int32, pin_memory=True)
        self.oc_host = torch.empty(B, dtype=torch.int32, pin_memory=True)

    def _run_batch(self, Q, ks, efs):
        m = len(Q)
        self.q_host[:m].numpy()[:] = Q
        self.ke_host[0, :m].numpy()[:] = ks
        self.ke_host[1, :m].numpy()[:] = efs
        self.q_dev[:m].copy_(self.q_host[:m], non_blocking=True)
        self.k_dev[:m].copy_(self.ke_host[0, :m], non_blocking=True)
        self.ef_dev[:m].copy_(self.ke_host[1, :m], non_blocking=True)
        self.index.search_device(self.q_dev[:m], 0, self.max_ef, self.oi[:m], self.od[:m],
                                 self.oc[:m], store=self.store, rerank=self.rerank,
                                 summation=self.summation, k_per_query=self.k_dev[:m],
                                 ef_per_query=self.ef_dev[:m], async_=True)
        self.oi_host[:m].copy_(self.oi[:m], non_blocking=True)
        self.oc_host[:m].copy_(self.oc[:m], non_blocking=True)
        self.index.base.search_status()  # synchronises; raises on a kernel status
        return self.oi_host[:m].numpy(), self.oc_host[:m].numpy()

    def search_all(self, Q, ks, efs):
        """Closed loop: every request at once, in batches of max_batch."""
        ids = np.zeros((len(Q), self.max_ef), np.uint32)
        cnt = np.zeros(len(Q), np.int32)
        for s in range(0, len(Q), self.max_batch):
            e = min(len(Q), s + self.max_batch)
            oi, oc = self._run_batch(Q[s:e], ks[s:e], efs[s:e])
            ids[s:e] = oi.view(np.uint32)
            cnt[s:e] = oc
        return ids, cnt

    def serve(self, Q, ks, efs, arrivals_s) -> ServeResult:
        """Open loop: request i arrives arrivals_s[i] seconds after the start
        (non-decreasing)."""
        n = len(Q)
        lat = np.zeros(n)
        i = 0
        nb = 0
        t0 = time.perf_counter()
        while i < n:
            now = time.perf_counter() - t0
            if arrivals_s[i] > now:  # idle until the next arrival
                time.sleep(max(0.0, arrivals_s[i] - now - 50e-6))
                while time.perf_counter() - t0 < arrivals_s[i]:
                    pass
                now = time.perf_counter() - t0
            j = int(np.searchsorted(arrivals_s, now, side="right"))
            j = min(j, i + self.max_batch)
            self._run_batch(Q[i:j], ks[i:j], efs[i:j])
            done = time.perf_counter() - t0
            lat[i:j] = done - arrivals_s[i:j]
            nb += 1
            i = j
        wall = time.perf_counter() - t0
        return ServeResult(lat, nb, n / max(nb, 1), wall)


def poisson_arrivals(n: int, rate_qps: float, seed: int = 0) -> np.ndarray:
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate_qps, n))
