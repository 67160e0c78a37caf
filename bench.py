"""Benchmark: QPS at recall@10 >= 0.95 on one MI355X, FP8 gather + FP32 rerank
vs plain FP32 vs the host CPU, on a named ParlayANN workload.

A step is one batched search (qsearchAll, beamSearch.h:599-646) of every
query of the workload, with queries, graph and points resident in HBM.  Each
leg is calibrated first (untimed): the smallest beam of the sweep whose
recall@10 against FP64 brute-force groundtruth is >= 0.95 is the operating
point (check_nn_recall.h:184-265 sweeps beams the same way).  Then W warmup
and K timed steps; value = the headline leg's queries/s.

The host-CPU leg runs the oracle's restatement of the reference search
(oracle/, one thread per host core) on the same graph and queries at the
plain-FP32 operating beam, timed like checkRecall (check_nn_recall.h:41-59:
a 5M-element random permutation to flush caches, wall time of the batched
search only, 5 repetitions, median).

Prints one JSON line on stdout; progress goes to stderr.
"""
from __future__ import annotations

import argparse
import copy
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# the reference's beam sweep tops out at 1000 (check_nn_recall.h:214-216); the operating
# point is the smallest beam of this list whose recall reaches the target, so the list
# steps by at most 1/8 of the beam above 100
BEAMS = [10, 12, 14, 16, 18, 20, 22, 24, 28, 32, 36, 40, 45, 50, 56, 64, 72, 80, 90, 100, 112,
         128, 144, 160, 180, 200, 224, 256, 288, 320, 352, 384, 416, 448, 512, 576, 640, 720,
         800, 900, 1000]

WORKLOADS = {
    # configs: plain FP32 L2 beam search, 100k x 128 Gaussian mixture, Vamana R=32 L=64 a=1.2
    "f32-l2-100k": dict(kind="f32", metric="l2", n=100_000, nq=10_000, d=128, R=32, L=64,
                        alpha=1.2, k=10, legs=["fp32-tree", "fp32-inorder", "bf16-rerank",
                                               "fp8-rerank"], headline="fp8-rerank"),
    # exact-integer path: 1M x 128 uint8, Vamana R=64 L=128 a=1.2, beams 10..200
    "u8-l2-1M": dict(kind="u8", metric="l2", n=1_000_000, nq=10_000, d=128, R=64, L=128,
                     alpha=1.2, k=10, legs=["u8"], headline="u8",
                     beams=[10, 20, 50, 100, 200]),
    # precision: 10M x 96 unit anisotropic Gaussian, inner product, Vamana R=64 L=128 a=1.0
    "aniso-ip-10M": dict(kind="aniso", metric="ip", n=10_000_000, nq=100_000, d=96, R=64,
                         L=128, alpha=1.0, k=10, legs=["fp32-tree", "fp32-inorder",
                                                       "bf16-rerank", "fp8-rerank",
                                                       "q8-rerank"],
                         headline="fp8-rerank", uneven=True),
    # serving: HNSW M=32 (layer-0 degree 64) efc=200 over 1M x 200 unit vectors
    # (500-centre mixture, queries shifted 0.5 sigma), IP, k in {10, 100}
    "hnsw-ip-1M": dict(kind="hnsw", metric="ip", n=1_000_000, nq=10_000, d=200, m=32, efc=200,
                       k=10, legs=["fp32-tree", "bf16-rerank", "fp8-rerank", "fp8r-rerank",
                                   "q8-rerank"],
                       headline="fp8r-rerank"),
}
EFS = [16, 18, 20, 22, 24, 28, 32, 36, 40, 48, 64, 80, 100, 128, 160, 200, 256, 320, 400]


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_data(w):
    from parlayann_amd import datagen
    if w["kind"] == "hnsw":
        from parlayann_amd.index import _normalize
        X, Q = datagen.workload_hnsw_ip(n=w["n"], nq=w["nq"], d=w["d"])
        return _normalize(X), _normalize(Q)
    if w["kind"] == "f32":
        return datagen.workload_f32_l2(n=w["n"], nq=w["nq"], d=w["d"])
    if w["kind"] == "u8":
        return datagen.workload_u8_l2(n=w["n"], nq=w["nq"], d=w["d"])
    if w["kind"] == "aniso":
        return datagen.workload_aniso_ip(n=w["n"], nq=w["nq"], d=w["d"])
    raise ValueError(w["kind"])


LEG_OPTS = {
    "fp32-tree": dict(store="native", summation="tree", rerank=False),
    "fp32-inorder": dict(store="native", summation="inorder", rerank=False),
    "bf16-rerank": dict(store="bf16", summation="tree", rerank=True),
    "fp8-rerank": dict(store="fp8", summation="tree", rerank=True),
    # FP8 of each point's residual to the nearest of 2048 sampled points (MIPS)
    "fp8r-rerank": dict(store="fp8r", summation="tree", rerank=True),
    # the reference's own 8-bit scalar quantization (its Python quant path)
    "q8-rerank": dict(store="q8", summation="tree", rerank=True),
    "u8": dict(store="native", summation="tree", rerank=False),
}


def row_bytes(leg, d, kind):
    if leg.startswith("bf16"):
        return 2 * d
    if leg.startswith("fp8r"):
        return d + 4  # the row's centroid id
    if leg.startswith("fp8") or leg.startswith("q8") or kind == "u8":
        return d
    return 4 * d


def max_rel_dist_error(X, Q, ids, dists, metric, chunk=8192):
    """max |d - d64| / |d64| over the returned items, d64 the FP64 distance."""
    worst = 0.0
    for s in range(0, len(Q), chunk):
        I = ids[s:s + chunk].astype(np.int64)
        P = X[I.reshape(-1)].astype(np.float64).reshape(I.shape + (X.shape[1],))
        q = Q[s:s + chunk].astype(np.float64)[:, None, :]
        d64 = ((P - q) ** 2).sum(-1) if metric == "l2" else -(P * q).sum(-1)
        err = np.abs(dists[s:s + chunk].astype(np.float64) - d64) / np.maximum(np.abs(d64), 1e-30)
        worst = max(worst, float(err.max()))
    return worst


def calib_copy(dev):
    """An elementwise kernel that reads 1 GiB and writes 1 GiB (its name
    contains "neg"): the scale of rocprofv3's FETCH_SIZE / WRITE_SIZE."""
    import torch
    a_ = torch.ones(1 << 28, dtype=torch.float32, device=dev)
    b_ = torch.empty_like(a_)
    for _ in range(3):
        torch.neg(a_, out=b_)
    torch.cuda.synchronize()
    del a_, b_


def gather_ceiling(dev):
    """The device's random row-gather ceiling (parlayann_amd/libpann_calib.so):
    2^26 8-lane reads of the first read_bytes of pseudo-random 128-byte rows of
    a 4 GiB buffer.  Every read fetches its whole 128-byte line, so the line
    rate, not HBM's streaming bandwidth, bounds the search's gathers."""
    import ctypes
    import torch
    lib = ctypes.CDLL(os.path.join(ROOT, "parlayann_amd", "libpann_calib.so"))
    fn = lib.pann_calib_gather
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32,
                   ctypes.c_uint64, ctypes.c_void_p, ctypes.POINTER(ctypes.c_float),
                   ctypes.c_void_p]
    nb, nreads = 4 << 30, 1 << 26
    buf = torch.zeros(nb, dtype=torch.uint8, device=dev)
    sink = torch.zeros(16, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    torch.cuda.synchronize()
    out = {}
    # (row stride, bytes read): whole lines, part lines, and 96-byte rows packed
    # at a 96-byte stride (half of them straddle two lines)
    for stride, rb in ((128, 128), (128, 96), (128, 32), (96, 96)):
        ms = []
        for _ in range(3):
            t = ctypes.c_float()
            rc = fn(buf.data_ptr(), nb, stride, rb, nreads, sink.data_ptr(), ctypes.byref(t),
                    stream)
            if rc != 0:
                raise RuntimeError(f"pann_calib_gather failed: hipError {rc}")
            ms.append(t.value)
        best = min(ms)
        out[f"read_{rb}B_of_{stride}B_rows"] = {"ms": best, "rows_per_s": nreads / best * 1e3,
                                                "line_gbs": nreads * 128 / best / 1e6}
    del buf, sink
    return out


def measured_traffic(workload, w, leg, beam):
    """HBM bytes per launch of the search kernel for this exact workload, leg
    and beam, from the committed rocprofv3 FETCH_SIZE + WRITE_SIZE summary
    (profiles/traffic.json, each counter scaled by what it reads for a
    device copy of known size); None when no such measurement exists."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    if not os.path.exists(p):
        return None
    with open(p) as f:
        rows = json.load(f)
    for r in rows:
        if (r["workload"], r["n"], r["nq"], r["leg"], r["beam"]) == \
                (workload, w["n"], w["nq"], leg, beam):
            return r["traffic_bytes_per_launch"]
    return None


def parse_args(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="aniso-ip-10M", choices=sorted(WORKLOADS))
    ap.add_argument("--target-recall", type=float, default=0.95)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--legs", default="")
    ap.add_argument("--n", type=int, default=0, help="override n (smaller runs; not the metric)")
    ap.add_argument("--nq", type=int, default=0)
    ap.add_argument("--rerank-factor", type=float, default=100.0,
                    help="rerank min(k * factor, beam) entries (QueryParams default 100, "
                         "types.h:237: the whole beam for k=10, beam <= 1000)")
    ap.add_argument("--beam", type=int, default=0,
                    help="fixed operating beam (skips the recall sweep; used for profiling)")
    ap.add_argument("--no-uneven", action="store_true")
    ap.add_argument("--no-serve", action="store_true",
                    help="HNSW workload: skip the open-loop serving runs (profiling)")
    ap.add_argument("--calib-copy", action="store_true",
                    help="run a 1 GiB read + 1 GiB write elementwise kernel first (calibrates "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE against a copy of known size)")
    ap.add_argument("--no-calib-gather", action="store_true",
                    help="skip the random row-gather ceiling measurement (about 0.1 s)")
    ap.add_argument("--cpu-seconds", type=float, default=6.0,
                    help="target wall time of one CPU-baseline repetition")
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the launch-shape autotuner (library defaults)")
    ap.add_argument("--cache", default="",
                    help="directory holding the built graph and groundtruth of this workload "
                         "(written on the first run, loaded by later ones, e.g. profiler passes)")
    ap.add_argument("--no-secondary", action="store_true",
                    help="run only --workload (by default the 10M inner-product headline is "
                         "followed by the 100k x 128 FP32, 1M x 128 uint8 and HNSW 1M x 200 "
                         "workloads, reported under config.secondary)")
    return ap.parse_args(argv)


def cpu_model():
    """The host CPU's model name (/proc/cpuinfo), reported with every CPU number."""
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


_NATIVE_ORACLE = []


def cpu_oracle():
    """The oracle for the host-CPU leg, built with -march=native on this host (the
    reference's flags, parallelDefsANN:8) into oracle/liboracle_native.so; the
    portable build if that compile fails.  Returns (context manager, flags)."""
    import contextlib
    from oracle import pyoracle
    if not _NATIVE_ORACLE:
        path = os.path.join(ROOT, "oracle", "liboracle_native.so")
        try:
            subprocess.run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle"),
                            "liboracle_native.so"], check=True, capture_output=True, timeout=120)
            _NATIVE_ORACLE.append((path, "-O3 -march=native"))
        except Exception as e:  # noqa: BLE001 -- report and use the portable build
            log(f"[bench] -march=native oracle build failed ({e}); portable build used")
            _NATIVE_ORACLE.append((None, "-O3 -march=x86-64-v3"))
    path, flags = _NATIVE_ORACLE[0]
    return (pyoracle.use_library(path) if path else contextlib.nullcontext()), flags


def main():
    args = parse_args()
    if args.cache:  # the autotuner's table lives beside the graph / groundtruth cache
        os.environ.setdefault("PANN_TUNE_CACHE", os.path.join(args.cache, "tune.json"))
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not part "
                                   "of it (--gpus must be 1)"}))
        sys.exit(2)
    out = run_workload(args, args.workload)
    if args.workload == "aniso-ip-10M" and not args.no_secondary and not (args.n or args.nq
                                                                         or args.beam):
        out["config"]["secondary"] = run_secondary(args)
    print(json.dumps(out))


def run_workload(args, name, w=None):
    w = dict(WORKLOADS[name]) if w is None else w
    if args.n:
        w["n"] = args.n
    if args.nq:
        w["nq"] = args.nq
    if w["kind"] == "hnsw":
        return run_hnsw(args, w, name)
    return run_vamana(args, w, name)


def run_vamana(args, w, name):
    import torch
    from parlayann_amd import DeviceIndex, autotune, datagen
    from parlayann_amd.groundtruth import groundtruth_device
    from parlayann_amd.recall import recall_plain, recall_with_ties
    from parlayann_amd.vamana import build_vamana

    legs = args.legs.split(",") if args.legs else w["legs"]
    beams = [args.beam] if args.beam else w.get("beams", BEAMS)
    dev = torch.device("cuda", 0)
    t0 = time.time()
    X, Q = make_data(w)
    log(f"[bench] data {X.shape} {X.dtype} + {Q.shape} in {time.time() - t0:.1f}s")
    t0 = time.time()
    perm = datagen.insertion_orders(len(X), 1, 4242)
    last = [time.time()]

    def progress(p, bi, lo, hi):  # one line per ~10 s of build work
        if time.time() - last[0] > 10:
            last[0] = time.time()
            log(f"[bench] build pass {p}: {hi:,}/{len(X):,} points inserted "
                f"({time.time() - t0:.0f}s)")

    tag = f"{name}-n{w['n']}-q{w['nq']}"
    cg = os.path.join(args.cache, tag + "-graph.npy") if args.cache else ""
    cgt = os.path.join(args.cache, tag + "-gt.npy") if args.cache else ""
    t_build = None
    if cg and os.path.exists(cg):
        adj = np.load(cg)
        log(f"[bench] graph loaded from {cg}")
    else:
        adj = build_vamana(X, w["metric"], w["R"], w["L"], w["alpha"], perm, progress=progress)
        t_build = time.time() - t0
        log(f"[bench] GPU Vamana build R={w['R']} L={w['L']} a={w['alpha']}: {t_build:.1f}s, "
            f"avg degree {adj[:, 0].mean():.2f}")
        if cg:
            os.makedirs(args.cache, exist_ok=True)
            np.save(cg, adj)
    idx = DeviceIndex(X, adj, w["metric"])
    t0 = time.time()
    if cgt and os.path.exists(cgt):
        gt = np.load(cgt)
    else:
        # brute-force candidates on the matrix cores, exact FP64 re-score
        gt, _ = groundtruth_device(idx, X, Q, 100, w["metric"])
        log(f"[bench] FP64 groundtruth (k=100, MFMA candidates) in {time.time() - t0:.1f}s")
        if cgt:
            np.save(cgt, gt)

    gceil = None if args.no_calib_gather else gather_ceiling(dev)
    if gceil:
        g = gceil["read_128B_of_128B_rows"]
        log(f"[bench] random 128-byte row gather ceiling: {g['rows_per_s'] / 1e9:.1f} G rows/s "
            f"= {g['line_gbs']:,.0f} GB/s of lines")
    if args.calib_copy:
        calib_copy(dev)
    for leg in legs:
        st = LEG_OPTS[leg]["store"]
        if st != "native":
            idx.build_store(st)
    k = w["k"]
    nq = len(Q)
    tq = torch.from_numpy(np.ascontiguousarray(Q)).to(dev)
    oi = torch.zeros((nq, k), dtype=torch.int32, device=dev)
    od = torch.zeros((nq, k), dtype=torch.float32, device=dev)
    ov = torch.zeros(nq, dtype=torch.int32, device=dev)
    oc = torch.zeros(nq, dtype=torch.int32, device=dev)
    avg_deg = float(adj[:, 0].mean())
    results = {}
    for leg in legs:
        opts = LEG_OPTS[leg]
        chosen, best = None, None
        for B in beams:
            if B < k:
                continue
            idx.search_device(tq, k, B, oi, od, out_visited=ov, out_cmps=oc,
                              rerank_factor=args.rerank_factor, **opts)
            torch.cuda.synchronize()
            ids = oi.cpu().numpy().view(np.uint32)
            rec = recall_plain(gt, ids, k)
            log(f"[bench] {leg} beam {B}: recall@10 {rec:.4f}")
            if best is None or rec > best[1] + 1e-4:
                best = (B, rec, ids, od.cpu().numpy())
            if rec >= args.target_recall:
                chosen = (B, rec, ids, od.cpu().numpy())
                break
        reached = chosen is not None
        if not reached:
            # the graph cannot reach the target (e.g. points unreachable from the
            # start point): report the highest-recall operating point instead
            log(f"[bench] {leg}: recall {args.target_recall} not reached; "
                f"using the best beam {best[0]} (recall {best[1]:.4f})")
            chosen = best
        B, rec, ids, dists = chosen
        rec_ties = recall_with_ties(X, Q, gt, ids, k, w["metric"])
        mre = max_rel_dist_error(X, Q, ids, dists, w["metric"] if w["kind"] != "u8" else "l2")
        tuned = {"waves_per_cu": 0, "visits_per_round": 0, "eval_cache_log2": 0,
                 "frontier_filter": 0, "exact_counts": False, "hash_slots_log2": 0,
                 "waves_per_query": 1}
        if not args.no_tune:
            tuned = autotune.tune(idx, tq, k, B, rerank_factor=args.rerank_factor, **opts)
            log(f"[bench] {leg}: tuned waves/CU {tuned['waves_per_cu']}, waves/query "
                f"{tuned['waves_per_query']}, " +
                (f"exact seen set (LDS slots 2^{tuned['hash_slots_log2'] or 'auto'})"
                 if tuned["exact_counts"] else
                 f"visits/round {tuned['visits_per_round']}, eval cache "
                 f"2^{tuned['eval_cache_log2']}, frontier filter "
                 f"{('auto', 'on', 'off')[tuned['frontier_filter']]}") +
                (" (cached)" if tuned["cached"] else "") +
                f" [{autotune.cache_path()} {tuned['key']}]")
        shape = {k_: tuned[k_] for k_ in ("waves_per_cu", "visits_per_round", "eval_cache_log2",
                                          "frontier_filter", "exact_counts", "hash_slots_log2",
                                          "waves_per_query")}
        for _ in range(args.warmup):
            idx.search_device(tq, k, B, oi, od, out_visited=ov, out_cmps=oc, async_=True,
                              rerank_factor=args.rerank_factor, **opts, **shape)
        torch.cuda.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            idx.search_device(tq, k, B, oi, od, out_visited=ov, out_cmps=oc, async_=True,
                              rerank_factor=args.rerank_factor, **opts, **shape)
            kms.append(idx.last_kernel_ms())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        idx.search_status()
        qps = nq * args.steps / dt
        evals = oc.cpu().numpy().astype(np.float64)
        # one untimed launch with the reference's seen set gives full_dist_cmps
        # (beamSearch.h:178) per query: the algorithmic gather count; the timed
        # kernel re-evaluates ids it rejected earlier instead of remembering them
        idx.search_device(tq, k, B, oi, od, out_visited=ov, out_cmps=oc, exact_counts=True,
                          hash_slots_log2=shape["hash_slots_log2"],
                          rerank_factor=args.rerank_factor, **opts)
        torch.cuda.synchronize()
        vis = ov.cpu().numpy().astype(np.float64)
        cmps = oc.cpu().numpy().astype(np.float64)
        rb = row_bytes(leg, w["d"], w["kind"])
        qbytes = w["d"] * (1 if w["kind"] == "u8" else 4)
        nrr = min(k * args.rerank_factor, B) if opts["rerank"] else 0
        per_q = (vis * 4 * (1 + avg_deg) + cmps * rb + nrr * w["d"] * 4 + qbytes + k * 8)
        bytes_launch = float(per_q.sum())
        kms_avg = float(np.mean(kms))
        gbs = bytes_launch / (kms_avg * 1e-3) / 1e9
        results[leg] = dict(beam=B, recall_at_10=round(rec, 5), recall_ties=round(rec_ties, 5),
                            target_reached=reached, max_rel_dist_error=mre,
                            qps=qps, ms_per_step=dt * 1e3 / args.steps, kernel_ms=kms_avg,
                            avg_visited=float(vis.mean()), avg_dist_cmps=float(cmps.mean()),
                            avg_evals=float(evals.mean()),
                            bytes_per_launch=bytes_launch, achieved_gbs=gbs,
                            launch_shape=dict(shape, tuned=not args.no_tune))
        log(f"[bench] {leg}: beam {B} recall {rec:.4f} QPS {qps:,.0f} kernel {kms_avg:.3f} ms "
            f"{gbs:,.0f} GB/s visited {vis.mean():.1f} cmps {cmps.mean():.1f} "
            f"evals {evals.mean():.1f}")

    # the full beam sweep of the first leg: QPS and recall at every beam, and a
    # bit-exact check of a query sample against the oracle at each
    if w.get("sweep"):
        from oracle import pyoracle
        leg = legs[0]
        opts = LEG_OPTS[leg]
        shape = {k_: v for k_, v in results[leg]["launch_shape"].items() if k_ != "tuned"}
        m_chk = min(nq, 500)
        sweep = []
        for B in beams:
            if B < k:
                continue
            run = lambda a=True: idx.search_device(
                tq, k, B, oi, od, out_visited=ov, out_cmps=oc, async_=a,
                rerank_factor=args.rerank_factor, **opts, **shape)
            run(False)
            ids = oi.cpu().numpy().view(np.uint32)
            dists = od.cpu().numpy()
            vis = ov.cpu().numpy().view(np.uint32)
            o = pyoracle.search_batch(adj, X, w["metric"], Q[:m_chk], k, B)
            exact = bool(np.array_equal(ids[:m_chk], o["ids"]) and
                         np.array_equal(dists[:m_chk], o["dists"]) and
                         np.array_equal(vis[:m_chk], o["visited"]))
            rec = recall_plain(gt, ids, k)
            for _ in range(args.warmup):
                run()
            torch.cuda.synchronize()
            kms = []
            t0 = time.perf_counter()
            for _ in range(args.steps):
                run()
                kms.append(idx.last_kernel_ms())
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            idx.search_status()
            sweep.append(dict(beam=B, recall_at_10=round(rec, 5), qps=nq * args.steps / dt,
                              kernel_ms=float(np.mean(kms)), oracle_sample=m_chk,
                              bit_exact_vs_oracle=exact))
            log(f"[bench] {leg} sweep beam {B}: recall {rec:.4f} QPS {nq * args.steps / dt:,.0f}"
                f" oracle sample of {m_chk} bit-exact: {exact}")
        results[leg]["beam_sweep"] = sweep

    # batched uneven search: per-query k in {1, 10, 100}, beam ~ U[max(16, k), 256],
    # the whole query set in one persistent launch
    if w.get("uneven") and not args.no_uneven:
        rng = np.random.default_rng(77)
        ks = rng.choice([1, 10, 100], nq).astype(np.int32)
        bs = np.array([rng.integers(max(16, kk), 257) for kk in ks], np.int32)
        tk, tb = torch.from_numpy(ks).to(dev), torch.from_numpy(bs).to(dev)
        ui = torch.zeros((nq, 100), dtype=torch.int32, device=dev)
        ud = torch.zeros((nq, 100), dtype=torch.float32, device=dev)
        # the whole query set in one persistent launch, captured once into a
        # hipGraph and replayed each step
        gs = torch.cuda.Stream()
        for leg in [l for l in ("fp32-tree", "fp8-rerank") if l in legs]:
            opts = LEG_OPTS[leg]
            shape = {"waves_per_cu": 0, "visits_per_round": 0, "eval_cache_log2": 0,
                     "frontier_filter": 0, "waves_per_query": 1}
            if not args.no_tune:
                # replayed from a captured hipGraph: the default seen set
                t = autotune.tune(idx, tq, 100, 256, k_per_query=tk, beam_per_query=tb,
                                  out_stride=100, rerank_factor=args.rerank_factor,
                                  allow_exact=False, **opts)
                shape = {k_: t[k_] for k_ in ("waves_per_cu", "visits_per_round",
                                              "eval_cache_log2", "frontier_filter",
                                              "waves_per_query")}
            run = lambda a=False, st=None: idx.search_device(
                tq, 100, 256, ui, ud, out_visited=ov, out_cmps=oc, k_per_query=tk,
                beam_per_query=tb, async_=a, stream=st, rerank_factor=args.rerank_factor,
                **opts, **shape)
            run()
            torch.cuda.synchronize()
            ids = ui.cpu().numpy().view(np.uint32)
            sel = ks >= 10
            rec = recall_plain(gt[sel], ids[sel][:, :10], 10)
            with torch.cuda.stream(gs):
                run(True, gs)  # first use of the stream, outside the capture
            gs.synchronize()
            idx.search_status(gs)
            graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(graph, stream=gs):
                run(True, gs)
            with torch.cuda.stream(gs):  # replays run on the capture stream
                for _ in range(args.warmup):
                    graph.replay()
            torch.cuda.synchronize()
            ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
                  for _ in range(args.steps)]
            t0 = time.perf_counter()
            with torch.cuda.stream(gs):
                for e0, e1 in ev:
                    e0.record(gs)
                    graph.replay()
                    e1.record(gs)
            torch.cuda.synchronize()
            dt = time.perf_counter() - t0
            idx.search_status(gs)
            gms = [e0.elapsed_time(e1) for e0, e1 in ev]
            same = np.array_equal(ui.cpu().numpy().view(np.uint32)[sel][:, :10], ids[sel][:, :10])
            del graph
            # algorithmic bytes as for the uniform legs, per query with its own
            # k and beam: the reference's counts from one untimed eager launch
            # with its exact seen set
            idx.search_device(tq, 100, 256, ui, ud, out_visited=ov, out_cmps=oc,
                              k_per_query=tk, beam_per_query=tb, exact_counts=True,
                              rerank_factor=args.rerank_factor, **opts)
            torch.cuda.synchronize()
            vis = ov.cpu().numpy().astype(np.float64)
            cmps = oc.cpu().numpy().astype(np.float64)
            rb = row_bytes(leg, w["d"], w["kind"])
            nrr = (np.minimum(ks * args.rerank_factor, bs) if opts["rerank"] else 0)
            per_q = vis * 4 * (1 + avg_deg) + cmps * rb + nrr * w["d"] * 4 + w["d"] * 4 + ks * 8
            gms_avg = float(np.mean(gms))
            gbs = float(per_q.sum()) / (gms_avg * 1e-3) / 1e9
            results["uneven-" + leg] = dict(
                qps=nq * args.steps / dt, ms_per_step=dt * 1e3 / args.steps,
                graph_ms=gms_avg, recall_at_10_k_ge_10=round(rec, 5),
                replay_equals_eager=bool(same), k_choices=[1, 10, 100], beam_range=[16, 256],
                launch="hipGraph replay of one persistent launch",
                avg_visited=float(vis.mean()), avg_dist_cmps=float(cmps.mean()),
                bytes_per_launch=float(per_q.sum()), achieved_gbs=gbs,
                launch_shape=dict(shape, tuned=not args.no_tune))
            log(f"[bench] uneven {leg}: QPS {nq * args.steps / dt:,.0f} recall@10(k>=10) {rec:.4f}"
                f" replay {gms_avg:.2f} ms {gbs:,.0f} GB/s")

    cpu = None
    if not args.no_cpu:
        from oracle import pyoracle
        base_leg = "fp32-inorder" if "fp32-inorder" in results else next(iter(results))
        B = results[base_leg]["beam"]
        threads = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
        threads = min(threads, 64)
        # bounded sample: size one repetition to about --cpu-seconds
        ctx, flags = cpu_oracle()
        with ctx:
            m = min(nq, 1000)
            t0 = time.perf_counter()
            pyoracle.search_batch(adj, X, w["metric"], Q[:m], k, B, nthreads=threads)
            rate = m / (time.perf_counter() - t0)
            m = int(min(nq, max(1000, rate * args.cpu_seconds)))
            times = []
            for rep in range(5):
                _ = np.random.default_rng(rep).permutation(5_000_000)  # cache flush
                t0 = time.perf_counter()
                pyoracle.search_batch(adj, X, w["metric"], Q[:m], k, B, nthreads=threads)
                times.append(time.perf_counter() - t0)
        med = float(np.median(times))
        cpu = {"value": m / med, "unit": "queries/s", "cores": threads, "kind": "port",
               "cpu_model": cpu_model(), "build": f"gcc {flags}",
               "sample": f"first {m} of the {nq} queries at beam {B} (the in-order FP32 "
                         f"operating point), 5 repetitions, median; oracle restatement of "
                         f"beamSearch.h"}
        log(f"[bench] host CPU ({threads} threads): {cpu['value']:,.0f} QPS")

    head = w["headline"] if w["headline"] in results else next(iter(results))
    h = results[head]
    out = {
        "metric": "QPS at recall@10=0.95, 1 MI355X: FP8+rerank vs plain FP32 vs host CPU",
        "value": h["qps"],
        "unit": "queries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": h["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp8-rerank": "fp8", "fp8r-rerank": "fp8", "bf16-rerank": "bf16", "u8": "u8",
                  "q8-rerank": "i8" if w["metric"] == "ip" else "u8"}.get(head, "f32"),
        "data": "synthetic (seeded generator; graph built on the GPU by this library)",
        "config": {"workload": name, "n": w["n"], "dim": w["d"], "nq": nq, "k": k,
                   "graph": f"vamana R={w['R']} L={w['L']} alpha={w['alpha']}",
                   "headline_leg": head, "legs": results, "graph_build_s": t_build,
                   "gather_ceiling": gceil},
        "roofline": {"bound": "hbm", "achieved": h["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": h["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(name, w, head, h["beam"])},
        "cpu_baseline": cpu,
    }
    idx.close()
    return out


def run_hnsw(args, w, name):
    """HNSW serving workload: calibrate ef per leg, closed-loop batches of all
    queries (inputs resident), then open-loop Poisson arrivals through the
    continuous-batching server at 50% and 80% of the closed-loop rate."""
    import torch
    from parlayann_amd import autotune, hnsw
    from parlayann_amd.groundtruth import groundtruth_device
    from parlayann_amd.recall import recall_plain
    from parlayann_amd.serve import BatchServer, poisson_arrivals

    legs = args.legs.split(",") if args.legs else w["legs"]
    dev = torch.device("cuda", 0)
    t0 = time.time()
    X, Q = make_data(w)
    log(f"[bench] data {X.shape} + {Q.shape} in {time.time() - t0:.1f}s")
    tag = f"{name}-n{w['n']}-q{w['nq']}"
    cm = os.path.join(args.cache, tag + ".hnsw") if args.cache else ""
    cgt = os.path.join(args.cache, tag + "-gt.npy") if args.cache else ""
    t_build = None
    if cm and os.path.exists(cm):
        model = hnsw.read_model(cm)
        log(f"[bench] model loaded from {cm}")
    else:
        t0 = time.time()
        last = [time.time()]

        def progress(bi, bb, be):
            if time.time() - last[0] > 10:
                last[0] = time.time()
                log(f"[bench] HNSW build: {be:,}/{len(X):,} nodes ({time.time() - t0:.0f}s)")

        model = hnsw.build_hnsw(X, w["metric"], w["m"], w["efc"], 1.0, seed=4242,
                                progress=progress)
        t_build = time.time() - t0
        log(f"[bench] GPU HNSW build m={w['m']} efc={w['efc']}: {t_build:.1f}s, "
            f"top level {model.top_level}, avg layer-0 degree {model.graphs[0][:, 0].mean():.2f}")
        if cm:
            os.makedirs(args.cache, exist_ok=True)
            hnsw.write_model(cm, model)
    idx = hnsw.HNSWIndex(X, model, w["metric"])
    if cgt and os.path.exists(cgt):
        gt = np.load(cgt)
    else:
        t0 = time.time()
        gt, _ = groundtruth_device(idx.base, X, Q, 100, w["metric"])
        log(f"[bench] FP64 groundtruth (k=100, MFMA candidates) in {time.time() - t0:.1f}s")
        if cgt:
            np.save(cgt, gt)
    gceil = None if args.no_calib_gather else gather_ceiling(dev)
    if args.calib_copy:
        calib_copy(dev)
    for leg in legs:
        if LEG_OPTS[leg]["store"] != "native":
            idx.build_store(LEG_OPTS[leg]["store"])
    nq = len(Q)
    rng = np.random.default_rng(77)
    ks = rng.choice([10, 100], nq).astype(np.int32)
    tq = torch.from_numpy(Q).to(dev)
    tk = torch.from_numpy(ks).to(dev)
    avg_deg = float(model.graphs[0][:, 0].mean())
    results = {}
    efs_list = [args.beam] if args.beam else EFS
    for leg in legs:
        opts = LEG_OPTS[leg]
        kw = dict(store=opts["store"], rerank=opts["rerank"], summation=opts["summation"])
        chosen = None
        for ef in efs_list:
            efs = np.maximum(ks, ef).astype(np.int32)
            mx = int(efs.max())
            oi = torch.zeros((nq, mx), dtype=torch.int32, device=dev)
            od = torch.zeros((nq, mx), dtype=torch.float32, device=dev)
            oc = torch.zeros(nq, dtype=torch.int32, device=dev)
            te = torch.from_numpy(efs).to(dev)
            idx.search_device(tq, 0, mx, oi, od, oc, k_per_query=tk, ef_per_query=te, **kw)
            ids = oi.cpu().numpy().view(np.uint32)[:, :10]
            rec = recall_plain(gt, ids, 10)
            log(f"[bench] {leg} ef {ef}: recall@10 {rec:.4f}")
            chosen = (ef, rec, efs, mx, oi, od, oc, te)
            if rec >= args.target_recall:
                break
        ef, rec, efs, mx, oi, od, oc, te = chosen
        ov = torch.zeros(nq, dtype=torch.int32, device=dev)
        om = torch.zeros(nq, dtype=torch.int32, device=dev)
        launch = {"exact_counts": False, "hash_slots_log2": 0, "eval_cache_log2": 0,
                  "waves_per_cu": 0, "visits_per_round": 0}
        if not args.no_tune:
            t = autotune.tune_hnsw(idx, tq, tk, te, mx, **kw)
            launch = {k_: t[k_] for k_ in launch}
            log(f"[bench] {leg}: tuned seen set {'exact' if t['exact_counts'] else 'default'}"
                f" (LDS slots 2^{t['hash_slots_log2'] or 'auto'}, eval cache "
                f"2^{t['eval_cache_log2']}), waves/CU "
                f"{t['waves_per_cu'] or 'auto'}, visits/round {t['visits_per_round'] or 'auto'}"
                + (" (cached)" if t["cached"] else ""))
        ex_t = launch["exact_counts"]
        shape = {k_: v for k_, v in launch.items() if k_ != "exact_counts"}
        run = lambda a=True, ex=ex_t: idx.search_device(
            tq, 0, mx, oi, od, oc, k_per_query=tk, ef_per_query=te, out_visited=ov,
            out_cmps=om, async_=a, exact_counts=ex, **kw, **shape)
        for _ in range(args.warmup):
            run()
        torch.cuda.synchronize()
        kms = []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run()
            kms.append(idx.base.last_kernel_ms())
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        idx.base.search_status()
        qps = nq * args.steps / dt
        evals = om.cpu().numpy().astype(np.float64)
        run(False, True)  # the reference's dist_cmps (count_cmps over every level)
        vis = ov.cpu().numpy().astype(np.float64)
        cmps = om.cpu().numpy().astype(np.float64)
        rb = row_bytes(leg, w["d"], "f32")
        nrr = efs.astype(np.float64) if opts["rerank"] else 0
        per_q = vis * 4 * (1 + avg_deg) + cmps * rb + nrr * w["d"] * 4 + w["d"] * 4 + ks * 8
        bytes_launch = float(per_q.sum())
        kms_avg = float(np.mean(kms))
        gbs = bytes_launch / (kms_avg * 1e-3) / 1e9
        # open loop: Poisson arrivals through the continuous-batching server
        serving = {}
        srv = None if args.no_serve else BatchServer(idx, mx, 2 << 30, launch=launch, **kw)
        if srv is not None:  # untimed: host buffers and clocks settle before the runs
            nw = int(min(20_000, max(2_000, 0.25 * qps)))
            srv.serve(Q[np.arange(nw) % nq], ks[np.arange(nw) % nq], efs[np.arange(nw) % nq],
                      poisson_arrivals(nw, 0.5 * qps, seed=4))
        for frac in (() if args.no_serve else (0.5, 0.8)):
            rate = frac * qps
            nreq = int(min(200_000, max(20_000, rate * 2.0)))
            sel = np.arange(nreq) % nq
            res = srv.serve(Q[sel], ks[sel], efs[sel], poisson_arrivals(nreq, rate, seed=5))
            serving[f"{int(frac * 100)}pct"] = dict(offered_qps=rate, requests=nreq,
                                                    **res.summary())
            log(f"[bench] {leg} serve at {rate:,.0f} q/s: p50 {res.summary()['p50_ms']:.3f} ms "
                f"p99 {res.summary()['p99_ms']:.3f} ms, mean batch {res.mean_batch:.0f}")
        results[leg] = dict(ef=ef, ef_rule="max(ef, k)", recall_at_10=round(rec, 5),
                            target_reached=rec >= args.target_recall, qps=qps,
                            ms_per_step=dt * 1e3 / args.steps, kernel_ms=kms_avg,
                            avg_visited=float(vis.mean()), avg_dist_cmps=float(cmps.mean()),
                            avg_evals=float(evals.mean()), bytes_per_launch=bytes_launch,
                            achieved_gbs=gbs, serving=serving,
                            launch_shape=dict(launch, tuned=not args.no_tune))
        log(f"[bench] {leg}: ef {ef} recall {rec:.4f} QPS {qps:,.0f} kernel {kms_avg:.3f} ms "
            f"{gbs:,.0f} GB/s visited {vis.mean():.1f} cmps {cmps.mean():.1f}")
    cpu = None
    if not args.no_cpu and "fp32-tree" in results:
        from oracle import pyoracle
        ef = results["fp32-tree"]["ef"]
        threads = min(64, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity")
                      else os.cpu_count())
        adj = [np.ascontiguousarray(g) for g in model.graphs]
        ent = np.asarray(model.entrance, np.uint32)

        def cpu_run(m):
            for kk in (10, 100):
                sel = np.nonzero(ks[:m] == kk)[0]
                pyoracle.hnsw_search(adj, X, w["metric"], Q[sel], kk, max(ef, kk), ent,
                                     nthreads=threads)
        ctx, flags = cpu_oracle()
        with ctx:
            m = min(nq, 500)
            t0 = time.perf_counter()
            cpu_run(m)
            rate = m / (time.perf_counter() - t0)
            m = int(min(nq, max(500, rate * args.cpu_seconds)))
            times = []
            for rep in range(5):
                _ = np.random.default_rng(rep).permutation(5_000_000)  # cache flush
                t0 = time.perf_counter()
                cpu_run(m)
                times.append(time.perf_counter() - t0)
        cpu = {"value": m / float(np.median(times)), "unit": "queries/s", "cores": threads,
               "kind": "port", "cpu_model": cpu_model(), "build": f"gcc {flags}",
               "sample": f"first {m} of the {nq} queries (k in {{10,100}}, ef max(ef, k) at the "
                         f"FP32 operating ef {ef}), 5 repetitions, median; oracle restatement "
                         f"of HNSW.hpp search"}
        log(f"[bench] host CPU ({threads} threads): {cpu['value']:,.0f} QPS")
    # the headline is the fastest leg that reaches the recall target (FP8's three
    # mantissa bits cannot separate this mixture's nearest neighbours)
    ok = [l for l in results if results[l]["target_reached"]]
    head = max(ok, key=lambda l: results[l]["qps"]) if ok else next(iter(results))
    h = results[head]
    out = {
        "metric": "HNSW serving QPS at recall@10=0.95 (k in {10,100}), 1 MI355X",
        "value": h["qps"], "unit": "queries/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": h["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None,
        "dtype": {"fp8-rerank": "fp8", "bf16-rerank": "bf16",
                  "q8-rerank": "i8" if w["metric"] == "ip" else "u8"}.get(head, "f32"),
        "data": "synthetic (seeded generator; HNSW built on the GPU by this library)",
        "config": {"workload": name, "n": w["n"], "dim": w["d"], "nq": nq,
                   "k": "U{10,100}", "graph": f"hnsw m={w['m']} efc={w['efc']}",
                   "headline_leg": head, "legs": results, "graph_build_s": t_build,
                   "gather_ceiling": gceil},
        "roofline": {"bound": "hbm", "achieved": h["achieved_gbs"], "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": h["achieved_gbs"] / HBM_PEAK_GBS,
                     "traffic": measured_traffic(name, w, head, h["ef"])},
        "cpu_baseline": cpu,
    }
    idx.close()
    return out


# The workloads reported under config.secondary of the default run, with the
# legs each reports: the 100k x 128 FP32 L2 graph (recall saturates below 0.95,
# reported at its best beam), the exact-integer 1M x 128 uint8 beam sweep, and
# the HNSW serving workload (closed loop and Poisson arrivals, p50/p99).
SECONDARY = [
    ("f32-l2-100k", dict(legs=["fp32-tree", "fp32-inorder", "fp8-rerank"])),
    ("u8-l2-1M", dict(sweep=True)),
    ("hnsw-ip-1M", dict(legs=["fp32-tree", "fp8-rerank", "fp8r-rerank", "q8-rerank"])),
]


def run_secondary(args):
    import gc

    import torch
    out = {}
    for name, over in SECONDARY:
        a2 = copy.copy(args)
        a2.legs = ""
        a2.steps = min(args.steps, 10)
        a2.warmup = min(args.warmup, 2)
        a2.cpu_seconds = min(args.cpu_seconds, 3.0)
        a2.calib_copy = False
        a2.no_calib_gather = True
        w = dict(WORKLOADS[name])
        w.update(over)
        t0 = time.time()
        log(f"[bench] secondary workload {name}")
        r = run_workload(a2, name, w)
        r["config"].pop("gather_ceiling", None)
        r["config"]["wall_s"] = time.time() - t0
        for key in ("n_gpus", "higher_is_better", "scaling", "vs_baseline"):
            r.pop(key, None)
        out[name] = r
        gc.collect()
        torch.cuda.empty_cache()
    return out


if __name__ == "__main__":
    main()
