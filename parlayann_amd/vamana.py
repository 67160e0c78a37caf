"""Vamana graph construction on the device.

build_index / batch_insert (vamana/index.h:150-319) with every per-point step
on the GPU: the build-mode beam search that yields each insert's visited list
(:247-264), robustPrune over it (:267), update_neighbors (:271-273), and the
reverse-edge rule per grouped key (:292-303).  The host keeps the batch
schedule (:206-240) and the group_by_key of reverse edges (:281-285; values
kept in insertion order).  The insertion order is an input (the reference
draws parlay::random_permutation, :212).  Within a batch every search reads
the graph as it was before the batch, as in the reference, so the result
does not depend on scheduling.
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

from . import _lib as L
from .index import DeviceIndex


def batch_schedule(n: int, m: int, base: float = 2.0, max_fraction: float = 0.02):
    """[(floor, ceiling)] of batch_insert's prefix-doubling schedule
    (vamana/index.h:206-240)."""
    max_batch = int(max_fraction * float(np.float32(n)))
    max_batch = min(max_batch, 1000000)
    if max_batch == 0:
        max_batch = n
    out, count, inc = [], 0, 0
    while count < m:
        if math.pow(base, inc) <= max_batch:
            floor = int(math.pow(base, inc)) - 1
            ceiling = min(int(math.pow(base, inc + 1)) - 1, m)
            count = ceiling
        else:
            floor = count
            ceiling = min(count + max_batch, m)
            count += max_batch
        out.append((floor, ceiling))
        inc += 1
    return out


def _p(t):
    return t.data_ptr()


def build_vamana(points: np.ndarray, metric: str, R: int, L_build: int, alpha: float,
                 perms: np.ndarray, num_passes: int = 1, batch_factor: float = 0.125,
                 sort_neighbors: bool = True, device: int = 0, vis_cap: int | None = None,
                 index: DeviceIndex | None = None, progress=None,
                 timings: dict | None = None) -> np.ndarray:
    """Returns the n x (R + 1) adjacency (slot 0 = degree).

    timings: if a dict, the device is synchronised after each step and the
    seconds spent in each are added to it under "search" (build-mode beam
    search), "prune" (robustPrune of the visited lists), "update" (writing the
    new rows), "reverse_group" (grouping the reverse edges by target),
    "reverse_prune" (the reverse-edge rule and its prunes), "sort" (the final
    neighbour sort); "visited" counts the visited-list entries pruned,
    "reverse_edges" the grouped reverse edges and "<step>_evals" the row
    distances the search, prune and reverse_prune steps evaluated."""
    import time

    import torch

    n = points.shape[0]
    perms = np.asarray(perms, dtype=np.uint32).reshape(num_passes, n)
    idx = index or DeviceIndex(points, np.zeros((n, R + 1), np.uint32), metric, device)
    if idx.max_degree != R:
        raise ValueError("index max_degree must equal R")
    h = idx._h
    lib = L.lib()
    dev = torch.device("cuda", device)
    vis_cap = vis_cap or min(4096, max(8 * L_build, 512))
    i32 = torch.int32
    clock = [time.perf_counter()]

    def lap(name):  # phase timing (only when asked for: it synchronises)
        if timings is None:
            return
        torch.cuda.synchronize(dev)
        t = time.perf_counter()
        timings[name] = timings.get(name, 0.0) + (t - clock[0])
        if name in ("search", "prune", "reverse_prune"):  # row distances of the step
            ev = idx.build_eval_counts(reset=True)
            key = "search" if name == "search" else "prune"
            timings[name + "_evals"] = timings.get(name + "_evals", 0) + ev[key]
        clock[0] = time.perf_counter()

    if timings is not None:
        torch.cuda.synchronize(dev)
        idx.build_eval_counts(reset=True)
        clock[0] = time.perf_counter()
    for p in range(num_passes):
        a = alpha if p == num_passes - 1 else 1.0  # build_index :173-178
        order = torch.from_numpy(perms[p].astype(np.int32)).to(dev)
        for bi, (lo, hi) in enumerate(batch_schedule(n, n)):
            nb = hi - lo
            ids = order[lo:hi].contiguous()
            while True:
                vid = torch.empty((nb, vis_cap), dtype=i32, device=dev)
                vd = torch.empty((nb, vis_cap), dtype=torch.float32, device=dev)
                vc = torch.empty(nb, dtype=i32, device=dev)
                st = lib.pann_search_visited(h, _p(ids), nb, L_build, batch_factor, 0, vis_cap,
                                             _p(vid), _p(vd), _p(vc), None)
                if st == L.PANN_ERR_CAPACITY and vis_cap < 4096:
                    # a visited list outgrew the buffer: the batch only read the
                    # graph, so it is simply searched again with room to spare
                    vis_cap = min(4096, vis_cap * 2)
                    continue
                L.check(st)
                break
            lap("search")
            if timings is not None:
                timings["visited"] = timings.get("visited", 0) + int(vc.sum())
            out = torch.empty((nb, R), dtype=i32, device=dev)
            cnt = torch.empty(nb, dtype=i32, device=dev)
            L.check(lib.pann_robust_prune_batch(h, _p(ids), nb, _p(vid), _p(vd), _p(vc), vis_cap,
                                                a, R, 1, _p(out), _p(cnt), None))
            lap("prune")
            L.check(lib.pann_update_neighbors(h, _p(ids), nb, _p(out), _p(cnt), R, None))
            lap("update")
            # reverse edges: (ngh, index) pairs grouped by ngh (:281-285), on the
            # device: a stable sort by ngh keeps each group's values in
            # insertion (flattened) order
            mask = torch.arange(R, device=dev)[None, :] < cnt[:, None]
            keys = out[mask]
            if keys.numel() == 0:
                continue
            vals = ids[:, None].expand(nb, R)[mask]
            keys_s, order_k = torch.sort(keys, stable=True)
            vals_s = vals[order_k]
            ukeys, counts = torch.unique_consecutive(keys_s, return_counts=True)
            offs = torch.zeros(ukeys.numel() + 1, dtype=torch.int64, device=dev)
            offs[1:] = torch.cumsum(counts, 0)
            if timings is not None:
                timings["reverse_edges"] = timings.get("reverse_edges", 0) + int(keys.numel())
            lap("reverse_group")
            # groups by size class, one launch each: the launch's LDS (its sort
            # buffer, sized for its largest group) sets how many waves share a
            # CU, so a few large groups do not slow the many small ones; groups
            # past 1024 sort in global scratch
            classes = []
            lo_c = 0
            for hi_c in (64, 448, 1024, None):
                m = counts > lo_c if hi_c is None else (counts > lo_c) & (counts <= hi_c)
                classes.append(m.nonzero().flatten())
                lo_c = hi_c
            for sel in classes:
                ng = int(sel.numel())
                if ng == 0:
                    continue
                gk = ukeys[sel].to(i32).contiguous()
                gcnt = counts[sel]
                go = torch.zeros(ng + 1, dtype=torch.int64, device=dev)
                go[1:] = torch.cumsum(gcnt, 0)
                # gather the selected groups' values into one CSR array
                starts = offs[:-1][sel]
                rep_start = torch.repeat_interleave(starts - go[:-1], gcnt)
                gv = vals_s[torch.arange(int(go[-1]), device=dev) + rep_start].to(i32).contiguous()
                go32 = go.to(i32).contiguous()
                out2 = torch.empty((ng, R), dtype=i32, device=dev)
                cnt2 = torch.empty(ng, dtype=i32, device=dev)
                L.check(lib.pann_robust_prune_groups(h, _p(gk), ng, _p(go32), _p(gv),
                                                     int(gcnt.max()), a, R, 2, _p(out2),
                                                     _p(cnt2), None))
                L.check(lib.pann_update_neighbors(h, _p(gk), ng, _p(out2), _p(cnt2), R, None))
            lap("reverse_prune")
            if progress:
                progress(p, bi, lo, hi)
    if sort_neighbors:  # :180-185, ties by id
        chunk = 1 << 18
        for lo in range(0, n, chunk):
            hi = min(n, lo + chunk)
            ids = torch.arange(lo, hi, dtype=i32, device=dev)
            zero = torch.zeros(hi - lo, dtype=i32, device=dev)
            out = torch.empty((hi - lo, R), dtype=i32, device=dev)
            cnt = torch.empty(hi - lo, dtype=i32, device=dev)
            L.check(lib.pann_robust_prune_batch(h, _p(ids), hi - lo, _p(zero), None, _p(zero), 0,
                                                1.0, R, 3, _p(out), _p(cnt), None))
            L.check(lib.pann_update_neighbors(h, _p(ids), hi - lo, _p(out), _p(cnt), R, None))
        lap("sort")
    adj = idx.download_graph()
    if index is None:
        idx.close()
    return adj


def _main(argv=None):
    """Builds the graph of the inner-product workload (unit-normalised
    anisotropic Gaussian, seeded) and prints one JSON line with the build time
    split into its steps; for profiling the build under rocprofv3."""
    import argparse
    import json
    import time

    import torch

    from . import datagen
    ap = argparse.ArgumentParser(prog="python -m parlayann_amd.vamana")
    ap.add_argument("--n", type=int, default=10_000_000)
    ap.add_argument("--d", type=int, default=96)
    ap.add_argument("--R", type=int, default=64)
    ap.add_argument("--L", type=int, default=128)
    ap.add_argument("--alpha", type=float, default=1.0)
    ap.add_argument("--split", action="store_true",
                    help="synchronise after each step and report its time")
    args = ap.parse_args(argv)
    X, _ = datagen.workload_aniso_ip(n=args.n, nq=1, d=args.d)
    perm = datagen.insertion_orders(len(X), 1, 4242)
    print(f"[build] aniso-ip n={args.n} d={args.d} R={args.R} L={args.L} alpha={args.alpha}",
          flush=True)
    idx = DeviceIndex(X, np.zeros((len(X), args.R + 1), np.uint32), "ip")
    torch.cuda.synchronize()
    tm = {} if args.split else None
    t0 = time.perf_counter()
    adj = build_vamana(X, "ip", args.R, args.L, args.alpha, perm, index=idx, timings=tm,
                       progress=lambda p, bi, lo, hi: print(f"[build] batch {bi} [{lo}, {hi})",
                                                            flush=True) if bi % 8 == 0 else None)
    dt = time.perf_counter() - t0
    idx.close()
    print(json.dumps(dict(n=args.n, d=args.d, R=args.R, L=args.L, alpha=args.alpha,
                          seconds=dt, avg_degree=float(adj[:, 0].mean()), split=tm)), flush=True)


if __name__ == "__main__":
    _main()
