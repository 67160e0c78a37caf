"""Launch-shape autotuner for the batched beam search.

The search's results do not depend on its launch shape, only its speed does.
For an index, a store and a beam (or a per-query k / beam mix) the tuner
times every (waves per CU, visits gathered per round, evaluated-id cache
size) choice on a sample of the caller's queries -- the persistent grid's
occupancy cap, whether the next visit's rows are gathered together with the
current one's while a merge is deferred, and how many LDS slots remember
evaluated ids (fewer rows gathered again, against fewer waves per CU) -- and
keeps the four fastest.  The reference's exact seen set is timed on the
sample too, at a few LDS sizes and occupancies (an id seen once is never
evaluated again, at the cost of an insert per neighbour and, past LDS, a
global table).  Workgroups of 2 and 4 waves per query
(each visit's gathers, the merges and the rerank shared) are timed too.
The four, the best exact shape, and the frontier-id filter
forced on and off for the best (fewer rows gathered again, against a filter
test before every gather) are then timed on the whole query batch, whose
throughput a small sample's tail can misstate.  Winners are cached in a JSON file keyed by
device, index shape, store and beam, so later runs reuse them.
"""
from __future__ import annotations

import json
import os
from typing import Optional

WAVES_PER_CU = (4, 6, 8, 12, 16)
VISITS_PER_ROUND = (1, 2, 3)
EVAL_CACHE_LOG2 = (0, 8, 10, 12)
FRONTIER_FILTER = (0, 1, 2)  # automatic, on, off
# the reference's exact seen set (an id seen once is not evaluated again):
# (LDS slots log2 (0 = automatic), waves per CU); slower where the set
# overflows LDS, much faster where the default set re-evaluates many rows
EXACT_SHAPES = ((0, 0), (12, 0), (13, 0), (0, 8), (12, 8), (13, 8))
# several waves per query (a workgroup shares each visit's gathers, the merges
# and the rerank; no frontier filter): (waves per query, waves per CU)
COOP_SHAPES = ((2, 16), (2, 32), (4, 16), (4, 32))


def cache_path() -> str:
    return os.environ.get("PANN_TUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "parlayann_amd", "tune.json")


def _load() -> dict:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(table: dict):
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp"
        with open(tmp, "w") as f:
            json.dump(table, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # a read-only home: tuning still applies to this process


def tune(index, queries, k: int, beam: int, *, store: str = "native",
         summation: str = "tree", rerank: bool = False, k_per_query=None,
         beam_per_query=None, out_stride: Optional[int] = None, sample: int = 16384,
         reps: int = 2, use_cache: bool = True, rerank_factor: float = 100.0,
         confirm: int = 4, confirm_max: int = 131072, allow_exact: bool = True) -> dict:
    """Fastest {waves_per_cu, visits_per_round, eval_cache_log2,
    frontier_filter, exact_counts, hash_slots_log2, waves_per_query} for DeviceIndex `index`
    on device queries (a torch tensor).  allow_exact=False keeps the default
    seen set (a search captured into a hipGraph needs it).  Returns the
    choice and the timings."""
    import torch

    dev = queries.device
    uneven = k_per_query is not None or beam_per_query is not None
    from .index import device_info  # HIP's own figures (device names can be blank)
    di = device_info(dev.index or 0)
    key = "|".join(str(x) for x in (
        f"cus{di['num_cus']}-hbm{di['hbm_bytes']}", index.n, index.dim, index.max_degree,
        str(index.dtype), index.metric, store, summation, int(rerank), k, beam,
        "uneven" if uneven else "uniform", "exact" if allow_exact else "noexact"))
    table = _load() if use_cache else {}
    if use_cache and key in table and "waves_per_query" in table[key]:
        return dict(table[key], cached=True, key=key)
    stride = out_stride or k
    mfull = min(int(queries.shape[0]), confirm_max)
    oi = torch.empty((mfull, max(stride, 1)), dtype=torch.int32, device=dev)
    od = torch.empty((mfull, max(stride, 1)), dtype=torch.float32, device=dev)

    def time_shape(m, wpc, vpr, ec, ff, ex=False, hs=0, wpq=1):
        kw = dict(store=store, summation=summation, rerank=rerank, waves_per_cu=wpc,
                  visits_per_round=vpr, eval_cache_log2=ec, frontier_filter=ff,
                  exact_counts=ex, hash_slots_log2=hs, waves_per_query=wpq,
                  k_per_query=k_per_query[:m] if k_per_query is not None else None,
                  beam_per_query=beam_per_query[:m] if beam_per_query is not None else None,
                  rerank_factor=rerank_factor)
        index.search_device(queries[:m], k, beam, oi[:m], od[:m], **kw)  # warm
        ms = []
        for _ in range(reps):
            index.search_device(queries[:m], k, beam, oi[:m], od[:m], async_=True, **kw)
            ms.append(index.last_kernel_ms())
        index.search_status()
        return min(ms)

    m = min(mfull, sample)
    timings = []
    for wpc in WAVES_PER_CU:
        for vpr in VISITS_PER_ROUND:
            for ec in EVAL_CACHE_LOG2:
                timings.append((time_shape(m, wpc, vpr, ec, 0), wpc, vpr, ec, 0, False, 0, 1))
    for wpq, wpc in COOP_SHAPES:
        for vpr in VISITS_PER_ROUND:
            # the cooperative launch takes no frontier filter (2 = off)
            timings.append((time_shape(m, wpc, vpr, 0, 2, wpq=wpq), wpc, vpr, 0, 2, False, 0,
                            wpq))
    timings.sort()
    shapes = [t[1:] for t in timings[:confirm]]
    bw, bv, be = next((t[1], t[2], t[3]) for t in timings if t[7] == 1)
    shapes += [(bw, bv, be, ff, False, 0, 1) for ff in FRONTIER_FILTER[1:]]
    if allow_exact:
        ex_t = sorted((time_shape(m, wpc, 1, 0, 0, True, hs), wpc, 1, 0, 0, True, hs, 1)
                      for hs, wpc in EXACT_SHAPES)
        timings += ex_t
        shapes.append(ex_t[0][1:])
    full = [(time_shape(mfull, *sh), *sh) for sh in shapes]
    best = min(full)
    res = {"waves_per_cu": best[1], "visits_per_round": best[2], "eval_cache_log2": best[3],
           "frontier_filter": best[4], "exact_counts": best[5], "hash_slots_log2": best[6],
           "waves_per_query": best[7], "ms": best[0], "queries": mfull, "sample": m,
           "timings": [{"waves_per_cu": w, "visits_per_round": v, "eval_cache_log2": e,
                        "frontier_filter": f_, "exact_counts": x, "hash_slots_log2": h,
                        "waves_per_query": q_, "ms": t, "queries": n_}
                       for n_, rows in ((m, timings), (mfull, full))
                       for t, w, v, e, f_, x, h, q_ in rows]}
    if use_cache:
        table[key] = res
        _store(table)
    return dict(res, cached=False, key=key)


# (exact seen set, its LDS slots log2, evaluated-id cache log2 of the default set)
HNSW_SHAPES = ((False, 0, 0), (False, 0, 10), (False, 0, 11), (True, 0, 0), (True, 12, 0))
# (a serving batch's launch shares the device with the next one's, so the
# HNSW shapes leave room: 16 waves per CU times a batch no faster than 12)
HNSW_WAVES = (12, 10, 8)
HNSW_VISITS_PER_ROUND = (2, 3)


def tune_hnsw(index, queries, k_per_query, ef_per_query, max_ef: int, *, store: str = "native",
              summation: str = "tree", rerank: bool = False, reps: int = 3,
              use_cache: bool = True) -> dict:
    """Fastest seen set and occupancy for HNSWIndex `index` on a batch of
    device queries with per-query k and ef: the default set (every
    neighbour evaluated, checked against the frontier and pending
    candidates), with or without an LDS cache of evaluated ids, or the
    reference's exact set, which at small ef and high degree gathers far
    fewer rows (an id already seen is not evaluated again); results are the
    same.  Returns {exact_counts, hash_slots_log2, eval_cache_log2,
    waves_per_cu, visits_per_round, ms}."""
    import torch
    from .index import device_info
    dev = queries.device
    di = device_info(dev.index or 0)
    base = index.base
    key = "|".join(str(x) for x in (
        "hnsw", f"cus{di['num_cus']}-hbm{di['hbm_bytes']}", base.n, base.dim,
        index.model.m, index.model.top_level, str(base.dtype), base.metric, store, summation,
        int(rerank), int(queries.shape[0]), max_ef))
    table = _load() if use_cache else {}
    if use_cache and key in table and "visits_per_round" in table[key]:
        return dict(table[key], cached=True, key=key)
    nq = int(queries.shape[0])
    oi = torch.empty((nq, max_ef), dtype=torch.int32, device=dev)
    od = torch.empty((nq, max_ef), dtype=torch.float32, device=dev)
    oc = torch.empty(nq, dtype=torch.int32, device=dev)
    timings = []
    for ex, hs, ec in HNSW_SHAPES:
        for wpc in HNSW_WAVES:
            # the lookahead only applies to the default seen set
            for vpr in ((0,) if ex else HNSW_VISITS_PER_ROUND):
                kw = dict(store=store, summation=summation, rerank=rerank, exact_counts=ex,
                          hash_slots_log2=hs, eval_cache_log2=ec, waves_per_cu=wpc,
                          visits_per_round=vpr, k_per_query=k_per_query,
                          ef_per_query=ef_per_query)
                index.search_device(queries, 0, max_ef, oi, od, oc, **kw)  # warm
                ms = []
                for _ in range(reps):
                    index.search_device(queries, 0, max_ef, oi, od, oc, async_=True, **kw)
                    ms.append(base.last_kernel_ms())
                base.search_status()
                timings.append((min(ms), ex, hs, ec, wpc, vpr))
    best = min(timings)
    res = {"exact_counts": best[1], "hash_slots_log2": best[2], "eval_cache_log2": best[3],
           "waves_per_cu": best[4], "visits_per_round": best[5], "ms": best[0],
           "timings": [{"exact_counts": e, "hash_slots_log2": h, "eval_cache_log2": c,
                        "waves_per_cu": w, "visits_per_round": v, "ms": t}
                       for t, e, h, c, w, v in timings]}
    if use_cache:
        table[key] = res
        _store(table)
    return dict(res, cached=False, key=key)
