#!/usr/bin/env python3
"""Benchmark: connected components (parallel_cc path) on one MI355X, plus the
min-cut and serving figures of the same metric.

Headline workload: R-MAT graph, 2^24 vertices, 2^28 undirected edges, (a,b,c,d)
= (0.45, 0.22, 0.22, 0.11), uint32 ids, loops and duplicates removed, generated
in HBM before timing, held both as the edge list and as a symmetric CSR (the
CSR is built from the edge list once, before timing; its build time is
reported).  A step = one CC call from the HBM-resident input to canonical
(minimum-id) labels in HBM and the component count on the host:
cac_cc_csr_dev (Afforest on the CSR; --cc-input csr, the default) or cac_cc_dev
(the edge list; --cc-input edges).  Both are timed in every run and reported
under "cc_paths"; value = whole-job GTEPS = 2^28 edges / seconds per step of the
chosen input.

The "extras" object carries the min-cut solves/sec and served-latency parts of
the metric, each on its own workload:
  c1  square-root min cut, Erdos-Renyi G(1000, 0.05), unit weights, p = 0.95;
  c3  dense Karger-Stein, complete graph n = 8192, weights U[1, 42], log2(n)^2 trials;
  c4  4096 planted-cut graphs, n log-uniform in [256, 65536], batched in one call
      vs the same kernels one graph per call;
  c5  served Poisson stream, 70% CC / 30% min cut, G(n, d/n), n log-uniform in
      [10^3, 10^6], d uniform in [8, 64], weights U[1, 100];
  ks  karger_stein (sparse Karger-Stein, p = 0.90) on Watts-Strogatz graphs of
      the reference's ws_32k_16 / ws_96k_16 shape, beside its published times;
  sq  square_root (p = 0.95) on the same two graphs, beside the reference's
      published 1-core and 1440-core times;
  rm  square_root and karger_stein on R-MAT graphs of the reference's
      rmat_16k_{1024,4096} shape (2^14 vertices, 16.4M / 65.5M edges), beside
      its published times;
  ap  approx_cut (p = 0.95) on the two Watts-Strogatz graphs (the reference
      publishes no approx_cut time), beside the oracle on one core.

Prints ONE JSON line (rank 0) with the roofline of the dominant kernel and a
CPU baseline (the oracle's union-find on the same edge list on min(16, cores)
threads, with its single-thread figure beside it).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "comm-avoiding-cuts-cc_amd"))

HBM_PEAK_GBS = 8000.0
METRIC = "CC GTEPS + min-cut solves/sec at ≥0.95 success; p50/p99 served latency"


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--scale", type=int, default=24)
    ap.add_argument("--edge-log2", type=int, default=28)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--prof-steps", type=int, default=5, help="extra steps timed per kernel with HIP events")
    ap.add_argument("--cpu-baseline-seconds", type=float, default=30.0,
                    help="budget for the CPU baseline sample (0 disables)")
    ap.add_argument("--copy-probe-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (PMC byte-counter scaling)")
    ap.add_argument("--extras", default="c1,c3,c4,c5,ks,sq,rm,ap",
                    help="comma list of min-cut / serving workloads to add (c1,c3,c4,c5,ks,sq,rm,ap) or 'none'")
    ap.add_argument("--c3-trials", type=int, default=169, help="dense KS trials (ceil(log2(8192))^2)")
    ap.add_argument("--c4-graphs", type=int, default=4096)
    ap.add_argument("--c4-stride", type=int, default=16, help="one-graph-per-call baseline on every k-th graph")
    ap.add_argument("--c4-ref-graphs", type=int, default=256,
                    help="graphs solved under the reference's trial rule (SquareRootCut::numberOfTrials)")
    ap.add_argument("--c5-queries", type=int, default=200)
    ap.add_argument("--c5-rate", type=float, default=20.0, help="offered load, queries per second")
    ap.add_argument("--ks-sizes", default="32000,96000", help="Watts-Strogatz vertex counts of the ks leg")
    ap.add_argument("--sq-sizes", default="32000,96000", help="Watts-Strogatz vertex counts of the sq leg")
    ap.add_argument("--rm-degrees", default="1000,4000", help="edges per vertex of the rm leg (R-MAT 2^14 vertices)")
    ap.add_argument("--cc-input", choices=("csr", "edges"), default="csr",
                    help="input the headline CC step reads: the symmetric CSR (Afforest) or the edge list")
    return ap.parse_args()


def kernel_bytes(name: str, m: int, n: int) -> float:
    """Algorithmic HBM bytes per launch of each CC kernel."""
    return {
        "cc_init": 4.0 * n,                # write parent
        "cc_runs": 8.0 * m,                # read every edge once
        "cc_settle": 8.0 * n,              # read + write parent
        "cc_bitmap": 4.0 * n + n / 8.0,    # read parent, write bitmap
        "cc_hook": 8.0 * m,                # read every edge once
        "cc_compress": 8.0 * n,            # read + write labels
        "af_init": 4.0 * n,                # write parent
        "af_hubs": 0.0,                    # the low-id prefix ranges (2^16 vertices): negligible
        "af_round": 16.0 * n,              # per vertex: offsets 4, first neighbour 4, its parent 4,
                                           # own parent written 4
        "af_mode": 0.0,                    # only when n fits the hub prefixes: negligible
        "af_final": 8.0 * n,               # parent of every vertex read, rewritten to its root
        "af_fix": n / 8.0,                 # the bitmap of vertices outside the sampled component
    }.get(name, 0.0)


def time_kernels(ctx, step, reps: int):
    """Average per-launch-group HIP-event time of each named kernel over reps steps."""
    ctx.prof_reset()
    ctx.prof_enable(True)
    for _ in range(reps):
        step()
    recs = ctx.prof_records()
    ctx.prof_enable(False)
    ctx.prof_reset()
    per, cnt = {}, {}
    for name, ms in recs:
        per[name] = per.get(name, 0.0) + ms
        cnt[name] = cnt.get(name, 0) + 1
    return {k: per[k] / cnt[k] for k in per}, {k: per[k] / reps for k in per}


def cpu_baseline(edges_host: np.ndarray, n: int, budget_s: float):
    """The oracle's union-find on the same edge list, on the host's cores.

    The headline figure is the multithreaded restatement (orc_cc_mt: one edge
    chunk per thread, CAS hooking) on OMP_NUM_THREADS threads when that is set
    (the host's CPU share per GPU), else on every core nproc reports, over the
    whole edge list, repeated until about a third of the budget is spent.  The single-thread union-find is timed beside it on
    the longest prefix that fits a third of the budget.
    """
    import oracle  # the CPU restatement: checker and baseline only

    m = edges_host.shape[0]
    nproc = os.cpu_count() or 1
    threads = max(1, min(nproc, int(os.environ.get("OMP_NUM_THREADS") or nproc)))
    reps, total = 0, 0.0
    while True:
        t0 = time.perf_counter()
        _, nc_mt = oracle.cc_mt(edges_host, n, 0, threads)
        total += time.perf_counter() - t0
        reps += 1
        if total >= budget_s / 3 or total * (reps + 1) / reps > budget_s / 2:
            break
    probe = min(m, 1 << 22)
    t0 = time.perf_counter()
    oracle.cc(edges_host[:probe], n)
    est = (time.perf_counter() - t0) * m / probe
    sample = m if est <= budget_s / 3 else max(probe, int(m * budget_s / 3 / est))
    t0 = time.perf_counter()
    oracle.cc(edges_host[:sample], n)
    t1 = time.perf_counter() - t0
    what = "the whole" if sample == m else f"the first {sample} edges of the"
    return {
        "value": m * reps / total / 1e9,
        "unit": "GTEPS",
        "cores": threads,
        "nproc": nproc,
        "threads_rule": "OMP_NUM_THREADS when set (the host's CPU share per GPU), else nproc",
        "kind": "port",
        "sample": f"the whole {m}-edge R-MAT edge list (sorted by (u,v)), {reps} pass(es) of the oracle's "
                  f"union-find on {threads} threads (CAS hooking, min-id roots), {total:.1f} s; "
                  f"{nc_mt} components",
        "single_thread": {"value": sample / t1 / 1e9, "unit": "GTEPS", "cores": 1,
                          "sample": f"{what} {m}-edge list, one pass, {t1:.1f} s"},
    }


def copy_probe(ctx, mib: int):
    """A plain device-to-device copy kernel of known size, for scaling PMC byte counters."""
    nbytes = mib << 20
    a = ctx.alloc(nbytes)
    b = ctx.alloc(nbytes)
    for _ in range(3):
        ctx.copy_dev(b.ptr, a.ptr, nbytes)
    ctx.synchronize()
    a.free()
    b.free()
    print(f"# copy probe: 3 copies of {nbytes} bytes", file=sys.stderr, flush=True)


F32_95 = float(np.float32(0.95))  # the reference parses p with std::stof


def _note(msg: str):
    """A progress line on stderr (a long leg must keep printing)."""
    print(f"# {msg}", file=sys.stderr, flush=True)


def bench_c1(ctx, cac, cpu_budget_s: float):
    """Square-root min cut on ER(1000, 0.05), unit weights, seed 0 (redrawn until connected)."""
    edges, seed_used = cac.gen_er(1000, 0.05, 0)
    n = 1000
    ctx.mincut(edges, n, F32_95, 0)
    reps = 10
    t0 = time.perf_counter()
    for _ in range(reps):
        cut, T = ctx.mincut(edges, n, F32_95, 0)
    dt = (time.perf_counter() - t0) / reps
    deg = np.bincount(edges["u"], edges["w"], n) + np.bincount(edges["v"], edges["w"], n)
    out = {"workload": f"square_root p=0.95 on G(1000, 0.05), unit weights, seed {seed_used}, m={len(edges)}",
           "solves_per_s": 1.0 / dt, "ms_per_solve": dt * 1e3, "cut": cut, "trials": T,
           "t": cac.intermediate_size(n, len(edges)), "min_degree": int(deg.min())}
    if cpu_budget_s > 0:
        import oracle  # the CPU restatement: baseline only
        e3 = np.stack([edges["u"], edges["v"], edges["w"]], axis=1).astype(np.uint64)
        t = out["t"]
        t0 = time.perf_counter()
        k = 0
        while k < T and time.perf_counter() - t0 < cpu_budget_s:
            oracle.sqrt_trial(n, e3, t, 0, k)
            k += 1
        per_trial = (time.perf_counter() - t0) / k
        out["cpu_baseline"] = {"value": 1.0 / (per_trial * T), "unit": "solves/s", "cores": 1, "kind": "port",
                               "sample": f"{k} of the {T} trials timed single-threaded, scaled to one solve"}
    return out


def bench_c3(ctx, cac, trials: int, cpu_budget_s: float = 0.0):
    """Dense Karger-Stein on the complete graph n = 8192, weights U[1, 42]."""
    n = 8192
    d_adj = ctx.alloc(n * n * 8)
    ctx.gen_complete_dev(n, 0, 42, d_adj.ptr)
    ctx.ks_dense_dev(d_adj.ptr, n, 0, 1 << 20, trials)  # warm-up: same size, other trial ids
    ctx.synchronize()
    t0 = time.perf_counter()
    raw, cut = ctx.ks_dense_dev(d_adj.ptr, n, 0, 0, trials)
    dt = time.perf_counter() - t0
    deg_min = None
    # the bound every trial starts from: minimum weighted degree (row sums on the host, one row at a time)
    import ctypes
    for r in range(0, n, 512):
        blk = np.empty((min(512, n - r), n), dtype=np.int64)
        ctx._check(ctx.lib.cac_memcpy_d2h(ctx.handle, blk.ctypes.data_as(ctypes.c_void_p), d_adj.ptr + r * n * 8,
                                          blk.nbytes))
        s = int(blk.sum(axis=1).min())
        deg_min = s if deg_min is None else min(deg_min, s)
    base = None
    if cpu_budget_s > 0:
        import oracle  # the reference's own library (oracle/_ref): baseline only
        if oracle.ref_available():
            iu = np.triu_indices(n, 1)
            w = np.empty(len(iu[0]), dtype=np.uint64)
            pos = 0
            for r in range(n - 1):  # the upper triangle of the device matrix, row by row
                row = np.empty(n, dtype=np.int64)
                ctx._check(ctx.lib.cac_memcpy_d2h(ctx.handle, row.ctypes.data_as(ctypes.c_void_p),
                                                  d_adj.ptr + r * n * 8, row.nbytes))
                w[pos: pos + n - 1 - r] = row[r + 1:]
                pos += n - 1 - r
            e3 = np.stack([iu[0].astype(np.uint64), iu[1].astype(np.uint64), w], axis=1)
            del iu, w
            t0 = time.perf_counter()
            ref_cut = oracle.ref_ks_mincut(n, e3, 0.0, 0)  # p = 0: exactly one trial (dense: m >= n^2/32)
            per_trial = time.perf_counter() - t0
            del e3
            base = {"value": 1.0 / per_trial, "unit": "trials/s", "cores": 1, "kind": "reference",
                    "sample": f"one trial of the reference's comincut::minimum_cut (p=0) on the same complete "
                              f"graph, {per_trial:.1f} s on one core (cut {ref_cut}); x{trials} trials = "
                              f"{per_trial * trials:.0f} s per solve"}
    d_adj.free()
    bytes_per_trial = 22.88e9  # sum over recursion children of 8 (s^2 + t'^2) + 8 * 128^2 per leaf
    return {"workload": f"dense Karger-Stein, complete graph n={n}, weights U[1,42], {trials} trials, seed 0",
            "seconds_per_solve": dt, "trials_per_s": trials / dt, "cut": int(cut), "min_degree": deg_min,
            "raw_cuts_equal_bound": bool(np.all(raw == deg_min)),
            "roofline": {"bound": "hbm", "algorithmic_bytes_per_trial": bytes_per_trial,
                         "achieved_gbs": bytes_per_trial * trials / dt / 1e9, "peak": HBM_PEAK_GBS,
                         "frac": bytes_per_trial * trials / dt / 1e9 / HBM_PEAK_GBS},
            "cpu_baseline": base}


# The reference's karger_stein (p = 0.90, one core) on its Watts-Strogatz inputs:
# mean seconds over its runs, evaluation/data/e1/baselines.csv (ws_32k_16:
# n = 32000, m = 512000; ws_96k_16: n = 96000, m = 1536000).
PUBLISHED_KS_WS_S = {32000: 59.59, 96000: 224.9}


def bench_ks(ctx, cac, sizes, cpu_budget_s: float):
    """karger_stein (sparse recursion, p = 0.90) on Watts-Strogatz graphs shaped
    like the reference's ws_{32k,96k}_16 inputs: ring lattice with 16 neighbours
    a side, 30% rewired, weight 100."""
    out = {}
    for n in sizes:
        e = cac.gen_ws(n, 16, 0)
        ctx.karger_stein(e, n, 0.90, 1)  # warm-up: same graph, another seed (buffers reach full size)
        ctx.synchronize()
        t0 = time.perf_counter()
        cut, T = ctx.karger_stein(e, n, 0.90, 0)
        dt = time.perf_counter() - t0
        deg = np.bincount(e["u"], e["w"], n) + np.bincount(e["v"], e["w"], n)
        r = {"workload": f"karger_stein p=0.90 seed 0, Watts-Strogatz n={n}, 16 ring neighbours a side, "
                         f"beta 0.3, weight 100, m={len(e)}",
             "seconds": dt, "trials": T, "trials_per_s": T / dt, "cut": int(cut), "min_degree": int(deg.min()),
             "published_reference_seconds": PUBLISHED_KS_WS_S.get(n),
             "speedup_vs_published": PUBLISHED_KS_WS_S[n] / dt if n in PUBLISHED_KS_WS_S else None}
        if cpu_budget_s > 0 and n == min(sizes):
            import oracle  # the reference's own library (oracle/_ref): baseline only
            if oracle.ref_available():
                e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
                t0 = time.perf_counter()
                oracle.ref_ks_mincut(n, e3, 0.0, 0)  # p = 0: exactly one trial
                per_trial = time.perf_counter() - t0
                r["cpu_baseline"] = {"value": 1.0 / per_trial, "unit": "trials/s", "cores": 1, "kind": "reference",
                                     "sample": f"one trial of the reference's comincut::minimum_cut (p=0) on the "
                                               f"same graph, {per_trial:.1f} s; x{T} trials = "
                                               f"{per_trial * T:.0f} s"}
        out[f"ws_{n // 1000}k_16"] = r
    return out


# The reference's square_root (p = 0.95) on its Watts-Strogatz inputs, median
# seconds over its runs: evaluation/data/e1/32k.csv and 96k.csv (1 core and
# 1440 cores of Piz Daint's multicore partition).
PUBLISHED_SQ_WS_S = {32000: {"1_core": 298.3, "1440_cores": 0.5548}, 96000: {"1_core": 2818.0, "1440_cores": 3.845}}


def bench_sq(ctx, cac, sizes, cpu_budget_s: float):
    """square_root (sampling contraction + dense Karger-Stein per trial, p = 0.95)
    on Watts-Strogatz graphs shaped like the reference's ws_{32k,96k}_16."""
    out = {}
    for n in sizes:
        e = cac.gen_ws(n, 16, 0)
        m = len(e)
        ctx.mincut(e, n, F32_95, 1)  # warm-up: same graph, another seed
        ctx.synchronize()
        t0 = time.perf_counter()
        cut, T = ctx.mincut(e, n, F32_95, 0)
        dt = time.perf_counter() - t0
        deg = np.bincount(e["u"], e["w"], n) + np.bincount(e["v"], e["w"], n)
        pub = PUBLISHED_SQ_WS_S.get(n)
        r = {"workload": f"square_root p=0.95 seed 0, Watts-Strogatz n={n}, 16 ring neighbours a side, beta 0.3, "
                         f"weight 100, m={m}",
             "seconds": dt, "solves_per_s": 1.0 / dt, "trials": T, "t": cac.intermediate_size(n, m),
             "trials_per_s": T / dt, "cut": int(cut), "min_degree": int(deg.min()),
             "published_reference_seconds": pub,
             "speedup_vs_published_1_core": pub["1_core"] / dt if pub else None,
             "speedup_vs_published_1440_cores": pub["1440_cores"] / dt if pub else None}
        if cpu_budget_s > 0 and n == min(sizes):
            import oracle  # the CPU restatement: baseline only
            e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
            t = r["t"]
            t0 = time.perf_counter()
            k = 0
            while k < T and time.perf_counter() - t0 < cpu_budget_s:
                oracle.sqrt_trial(n, e3, t, 0, k)
                k += 1
            per_trial = (time.perf_counter() - t0) / k
            r["cpu_baseline"] = {"value": 1.0 / per_trial, "unit": "trials/s", "cores": 1, "kind": "port",
                                 "sample": f"{k} of the {T} trials timed single-threaded ({per_trial:.2f} s each); "
                                           f"x{T} trials = {per_trial * T:.0f} s"}
        out[f"ws_{n // 1000}k_16"] = r
    return out


def bench_ap(ctx, cac, sizes, cpu_budget_s: float):
    """approx_cut (binomial subsampling, CC over five disjoint copies per level,
    p = 0.95) on the Watts-Strogatz graphs of the sq leg.  The edge list is
    uploaded inside each call, as the approx_cut command line does."""
    out = {}
    for n in sizes:
        e = cac.gen_ws(n, 16, 0)
        m = len(e)
        ctx.approx_cut(e, n, F32_95, 1)  # warm-up: same graph, another seed
        ctx.synchronize()
        reps = []
        for _ in range(5):
            t0 = time.perf_counter()
            a, T = ctx.approx_cut(e, n, F32_95, 0)
            reps.append(time.perf_counter() - t0)
        dt = sorted(reps)[len(reps) // 2]
        r = {"workload": f"approx_cut p=0.95 seed 0, Watts-Strogatz n={n}, 16 ring neighbours a side, beta 0.3, "
                         f"weight 100, m={m}",
             "seconds": dt, "seconds_of_5": reps, "solves_per_s": 1.0 / dt, "trials": T, "approximation": a}
        if cpu_budget_s > 0 and n == min(sizes):
            import oracle  # the CPU restatement: baseline and check
            e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
            t0 = time.perf_counter()
            want = oracle.approx_cut(n, e3, F32_95, 0)
            cpu_s = time.perf_counter() - t0
            r["equals_oracle"] = want == (a, T)
            r["cpu_baseline"] = {"value": 1.0 / cpu_s, "unit": "solves/s", "cores": 1, "kind": "port",
                                 "sample": f"one whole solve on one core ({cpu_s:.2f} s)"}
        out[f"ws_{n // 1000}k_16"] = r
    return out


# The reference on its R-MAT inputs rmat_16k_{1024,2048,4096} (n = 16000,
# m = 16.4M / 32.8M / 65.5M, unit weights), median seconds:
# square_root at 144 and 1296 cores (evaluation/data/e4/data.csv), karger_stein
# on one core (evaluation/data/e4/baselines.csv).
PUBLISHED_RMAT16K_S = {1000: {"square_root_144_cores": 11.37, "square_root_1296_cores": 5.885, "karger_stein_1_core": 12.13},
                       2000: {"square_root_144_cores": 23.2, "square_root_1296_cores": 7.812, "karger_stein_1_core": 31.32},
                       4000: {"square_root_144_cores": 32.93, "square_root_1296_cores": 9.931, "karger_stein_1_core": 46.39}}


def bench_rm(ctx, cac, degrees):
    """square_root (p = 0.95) and karger_stein (p = 0.90) on R-MAT graphs shaped
    like the reference's rmat_16k_{1024,2048,4096}: 2^14 vertices, 16.384M x
    (1, 2, 4) edges, unit weights (the generator's (0.45, 0.22, 0.22, 0.11))."""
    out = {}
    scale = 14
    n = 1 << scale
    for deg in degrees:
        m = n * deg
        d = ctx.alloc(m * 8)
        ctx.gen_rmat_dev(scale, m, 0, cac.rmat_thresholds(0.45, 0.22, 0.22), d.ptr)
        ue = d.download(np.uint32, 2 * m).reshape(m, 2)
        d.free()
        e = np.empty(m, dtype=cac.EDGE)
        e["u"], e["v"], e["w"] = ue[:, 0], ue[:, 1], 1
        del ue
        r = {"workload": f"R-MAT 2^{scale} vertices, {m} edges, unit weights, seed 0"}
        ctx.mincut(e, n, F32_95, 1)  # warm-up: same graph, another seed
        t0 = time.perf_counter()
        cut, T = ctx.mincut(e, n, F32_95, 0)
        r["square_root"] = {"seconds": time.perf_counter() - t0, "trials": T, "t": cac.intermediate_size(n, m),
                            "cut": int(cut)}
        ctx.karger_stein(e, n, 0.90, 1)
        t0 = time.perf_counter()
        cut2, T2 = ctx.karger_stein(e, n, 0.90, 0)
        r["karger_stein"] = {"seconds": time.perf_counter() - t0, "trials": T2, "cut": int(cut2)}
        pub = PUBLISHED_RMAT16K_S.get(deg)
        if pub:
            r["published_reference_seconds"] = pub
            r["square_root_speedup_vs_1296_cores"] = pub["square_root_1296_cores"] / r["square_root"]["seconds"]
            r["karger_stein_speedup_vs_1_core"] = pub["karger_stein_1_core"] / r["karger_stein"]["seconds"]
        out[f"rmat_16k_x{deg}"] = r
        del e
    return out


def _c4_run(ctx, cac, descs, edges, planted, stride: int, seed: int = 0, warm: bool = True, tag: str = ""):
    """Batched solve of descs (trials as set in descs) vs every stride-th graph
    one per call; returns the figures and the batch's cuts."""
    buf = ctx.alloc(edges.nbytes)
    buf.upload(edges)
    if warm:
        ctx.mincut_batch_dev(buf.ptr, descs, F32_95, seed + 1)  # warm-up at full size (another seed)
    t0 = time.perf_counter()
    cuts, trials = ctx.mincut_batch_dev(buf.ptr, descs, F32_95, seed)
    dt = time.perf_counter() - t0
    _note(f"c4 {tag}: {len(descs)} graphs, {int(trials.sum())} trials batched in {dt:.2f} s")
    sub = np.arange(0, len(descs), stride)
    t0 = time.perf_counter()
    single = []
    for g in sub:
        single.append(ctx.mincut_batch_dev(buf.ptr, descs[g:g + 1], F32_95, seed)[0][0])
        if time.perf_counter() - t0 > 60:
            raise RuntimeError("one-graph-per-call sample over 60 s")
    dt1 = time.perf_counter() - t0
    _note(f"c4 {tag}: {len(sub)} graphs one per call in {dt1:.2f} s")
    buf.free()
    batched_rate, single_rate = len(descs) / dt, len(sub) / dt1
    # the one-graph-per-call time of the whole batch, estimated from the sample
    # in proportion to each graph's trials x edges
    work = descs["trials"].astype(np.float64) * descs["m"].astype(np.float64)
    single_total_s = dt1 * work.sum() / work[sub].sum()
    return {"graphs": int(len(descs)), "edges": int(len(edges)), "total_trials": int(trials.sum()),
            "solves_per_s": batched_rate, "seconds": dt,
            "one_graph_per_call_solves_per_s": single_rate,
            "one_graph_per_call_sample": f"every {stride}th graph ({len(sub)} graphs, {dt1:.2f} s)",
            "speedup": batched_rate / single_rate,
            "speedup_work_weighted": single_total_s / dt,
            "success_fraction": float((cuts == planted).mean()),
            "never_below_planted": bool(np.all(cuts >= planted)),
            "batched_equals_single": bool(np.array_equal(np.array(single), cuts[sub]))}, cuts


def _c4_cpu_baseline(cac, descs, edges, t, budget_s: float):
    """The oracle's square-root trials (single-threaded port) on a sample of the
    batch's graphs, scaled to the batch: seconds per trial by graph size class."""
    import oracle  # the CPU restatement: baseline only

    order = np.argsort(descs["n"])
    picks = order[np.linspace(0, len(order) - 1, 12).astype(int)]
    per_trial, done, t0 = [], 0, time.perf_counter()
    for g in picks:
        d = descs[g]
        e = edges[int(d["edge_offset"]): int(d["edge_offset"] + d["m"])]
        e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
        s0 = time.perf_counter()
        oracle.sqrt_trial(int(d["n"]), e3, int(t[g]), 0, 0)
        per_trial.append((float(d["m"]), time.perf_counter() - s0))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    # seconds per trial grow with the edge count: fit a line through the sample
    ms_, ts_ = np.array([a for a, _ in per_trial]), np.array([b for _, b in per_trial])
    k = float(np.dot(ms_, ts_) / np.dot(ms_, ms_))
    total = float(k * (descs["trials"].astype(np.float64) * descs["m"].astype(np.float64)).sum())
    return {"value": len(descs) / total, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"one trial of each of {done} graphs spread over the size range, "
                      f"{ts_.sum():.1f} s, timed single-threaded; seconds per trial fitted as "
                      f"{k * 1e9:.1f} ns x edges and scaled to the batch's {int(descs['trials'].sum())} trials"}


def bench_c4(ctx, cac, count: int, stride: int, ref_count: int, cpu_budget_s: float):
    """Batched square-root min cut over planted-cut graphs vs one graph per call,
    under two trial rules, on the standard and the hard planted family."""
    out = {}
    t0 = time.perf_counter()
    descs, edges, planted = cac.gen_planted(0, count, 0)
    gen_s = time.perf_counter() - t0
    n = descs["n"].astype(np.int64)
    t = np.array([cac.intermediate_size(int(a), int(b)) for a, b in zip(n, descs["m"])])
    light = np.maximum(2, [cac.ks_trials(int(x), F32_95) for x in t])
    fam = (f"n log-uniform in [256, 65536] (mean {n.mean():.0f}), halves G(n/2, 16/(n/2-1)) weights U[1,100], "
           f"c in [1,8] unit cross edges, {len(edges)} edges, generated in {gen_s:.1f} s")
    # 1. the light rule over the full batch (the C4 configuration)
    descs["trials"] = light
    _note(f"c4: {count} planted graphs generated in {gen_s:.1f} s")
    r, _ = _c4_run(ctx, cac, descs, edges, planted, stride, tag="light rule")
    r["workload"] = f"{count} planted-cut graphs, " + fam
    r["trial_rule"] = ("light: T_g = max(2, ks_trials(t_g, 0.95)), t_g = intermediate_size(n_g, m_g); "
                       "no success bound of its own")
    r["launch_shapes"] = ctx.tuning_file or "built-in rule (no autotuner table)"
    if cpu_budget_s > 0:
        r["cpu_baseline"] = _c4_cpu_baseline(cac, descs, edges, t, cpu_budget_s)
    out.update(r)
    # the hard family (exact planted cut within 10% of the minimum weighted degree) under the light rule
    hd, he, hp = cac.gen_planted_hard(0, count, 0)
    ht = np.array([cac.intermediate_size(int(a), int(b)) for a, b in zip(hd["n"], hd["m"])])
    hd["trials"] = np.maximum(2, [cac.ks_trials(int(x), F32_95) for x in ht])
    hr, _ = _c4_run(ctx, cac, hd, he, hp, stride, tag="hard family, light rule")
    out["hard_family_light_rule"] = {
        "workload": f"{count} hard planted graphs (same sizes): halves of 8 Hamiltonian cycles with cycle weights "
                    f"U[1,100], planted cut floor(0.9 x minimum weighted degree) on c in [1,8] cross edges",
        "trial_rule": out["trial_rule"], **{k: hr[k] for k in ("success_fraction", "never_below_planted",
                                                                "total_trials", "solves_per_s")}}
    del he
    # 2. the reference's rule (SquareRootCut::numberOfTrials) on the first ref_count graphs
    rd, rp = descs[:ref_count].copy(), planted[:ref_count]
    re_ = edges[: int(rd["edge_offset"][-1] + rd["m"][-1])]
    rd["trials"] = [cac.sqrt_trials(int(a), int(b), F32_95) for a, b in zip(rd["n"], rd["m"])]
    rr, _ = _c4_run(ctx, cac, rd, re_, rp, max(1, stride // 2), warm=False, tag="reference rule")
    rr["workload"] = f"the first {ref_count} graphs of the same batch"
    rr["trial_rule"] = "reference: SquareRootCut::numberOfTrials (src/SquareRootCut.cpp:76-83), success >= 0.95 by its bound"
    hard_count = max(1, ref_count // 4)  # the hard family's trials run several times longer
    hd2, he2, hp2 = cac.gen_planted_hard(0, hard_count, 0)
    hd2["trials"] = [cac.sqrt_trials(int(a), int(b), F32_95) for a, b in zip(hd2["n"], hd2["m"])]
    h2, _ = _c4_run(ctx, cac, hd2, he2, hp2, max(1, stride // 2), warm=False, tag="hard family, reference rule")
    rr["hard_family"] = {"graphs": hard_count,
                         **{k: h2[k] for k in ("success_fraction", "never_below_planted", "total_trials",
                                               "solves_per_s", "seconds", "speedup", "speedup_work_weighted")}}
    out["reference_rule"] = rr
    return out


def light_trials(n: int, m: int) -> int:
    """A served min cut's trial budget: the Karger-Stein trial count for p = 0.95
    of its contracted graph (t = min(intermediate_size, 4096)), at least 2."""
    import cac
    t = min(cac.intermediate_size(n, m), 4096)
    return max(2, cac.ks_trials(t, F32_95))


def bench_c5(cac, count: int, rate: float, cpu_budget_s: float = 0.0):
    """Served Poisson stream of CC (70%) and min-cut (30%) queries."""
    from concurrent.futures import ThreadPoolExecutor

    rng = np.random.default_rng(0)
    ns = np.exp(rng.uniform(np.log(1e3), np.log(1e6), size=count)).astype(np.int64)
    ds = rng.uniform(8, 64, size=count)
    kinds = np.where(rng.random(count) < 0.7, cac.QUERY_CC, cac.QUERY_MINCUT)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    with ThreadPoolExecutor(16) as ex:
        edges = list(ex.map(lambda i: cac.gen_gnp(int(ns[i]), float(ds[i]), 0, i, 100), range(count)))
    # min-cut queries: a trial budget (the Karger-Stein trial count of the
    # contracted graph, at most 4096 vertices) and a seed of their own
    budget_T = [light_trials(int(ns[i]), len(edges[i])) for i in range(count)]
    recs = [cac.Server.query(int(kinds[i]), edges[i], int(ns[i]), F32_95, seed=i, trials=budget_T[i])
            for i in range(count)]
    budget = 160 << 30
    _note(f"c5: {count} queries generated")
    with cac.Server(0, workers=4, hbm_budget=budget, max_batch=64) as srv:
        t0 = time.perf_counter()
        res = srv.replay(recs, arrivals)
        wall = time.perf_counter() - t0
        peak, groups = srv.stats()
    _note(f"c5: stream served in {wall:.1f} s")
    lat = res["done_s"] - res["submit_s"]
    out = {"workload": f"{count} queries, Poisson arrivals at {rate}/s, 70% CC / 30% min cut (p=0.95), "
                       f"G(n, d/n), n log-uniform in [1e3, 1e6], d uniform in [8, 64], weights U[1,100], "
                       f"{sum(len(e) for e in edges)} edges in total; host edge lists, upload included",
           "offered_qps": rate, "throughput_qps": count / wall, "all_ok": bool(np.all(res["status"] == 0)),
           "latency_ms": {"p50": float(np.percentile(lat, 50) * 1e3), "p99": float(np.percentile(lat, 99) * 1e3)},
           "hbm_budget_gib": budget / 2**30, "peak_admitted_gib": peak / 2**30, "launch_groups": groups,
           "mean_group_size": float(res["batch_size"].mean())}
    for k, name in ((cac.QUERY_CC, "cc"), (cac.QUERY_MINCUT, "mincut")):
        lk = lat[kinds == k]
        out[f"{name}_latency_ms"] = {"count": int(len(lk)), "p50": float(np.percentile(lk, 50) * 1e3),
                                     "p99": float(np.percentile(lk, 99) * 1e3)}
    # served min-cut throughput: every min-cut query of the stream submitted at
    # once (closed loop), grouped on 4 workers, vs the same queries one graph
    # per launch group on one worker
    mci = [i for i in range(count) if kinds[i] == cac.QUERY_MINCUT]
    mc = [recs[i] for i in mci]
    rates, rates_up, all_rates, all_rates_up = {}, {}, {}, {}
    for name, w, mb in (("grouped", 4, 64), ("one_graph_per_launch", 1, 1)):
        with cac.Server(0, workers=w, hbm_budget=budget, max_batch=mb) as srv:
            # device-resident graphs (registered before the clock starts), as the
            # configuration's served state is
            hs = [srv.register_graph(edges[i], int(ns[i])) for i in mci]
            resident = [cac.Server.query(cac.QUERY_MINCUT, None, 0, F32_95, seed=i, trials=budget_T[i], graph=h)
                        for i, h in zip(mci, hs)]
            srv.replay(resident, np.zeros(len(resident)))  # warm-up: buffers reach their size
            # groups form by timing, so a replay can need larger buffers than
            # the last one grew: three timed replays, the median reported
            reps, reps_up = [], []
            for _ in range(3):
                t0 = time.perf_counter()
                r = srv.replay(resident, np.zeros(len(resident)))
                reps.append(len(mc) / (time.perf_counter() - t0))
                assert np.all(r["status"] == 0)
            rates[name] = float(np.median(reps))
            all_rates[name] = reps
            # the same queries bringing their edges (host upload inside the clock)
            for _ in range(3):
                t0 = time.perf_counter()
                ru = srv.replay(mc, np.zeros(len(mc)))
                reps_up.append(len(mc) / (time.perf_counter() - t0))
                assert np.all(ru["status"] == 0) and np.array_equal(r["cut"], ru["cut"])
            rates_up[name] = float(np.median(reps_up))
            all_rates_up[name] = reps_up
            _note(f"c5: {len(mc)} min-cut queries {name}: {rates[name]:.1f} queries/s resident, "
                  f"{rates_up[name]:.1f} with upload")
            if name == "grouped":
                ref = r["cut"].copy()
            else:
                out["mincut_grouped_equals_single"] = bool(np.array_equal(ref, r["cut"]))
    out["mincut_closed_loop_queries"] = len(mc)
    out["mincut_closed_loop_qps"] = rates
    out["mincut_closed_loop_qps_each_replay"] = {"resident": all_rates, "with_upload": all_rates_up}
    out["mincut_speedup"] = rates["grouped"] / rates["one_graph_per_launch"]
    out["mincut_closed_loop_qps_with_upload"] = rates_up
    out["mincut_speedup_with_upload"] = rates_up["grouped"] / rates_up["one_graph_per_launch"]
    if cpu_budget_s > 0:
        import oracle  # the CPU restatement: baseline only
        t0, done = time.perf_counter(), 0
        for i in np.argsort(ns):  # smallest queries first, until the budget is spent
            e = edges[i]
            if kinds[i] == cac.QUERY_CC:
                oracle.cc(np.stack([e["u"], e["v"]], 1).astype(np.uint32), int(ns[i]))
            else:
                e3 = np.stack([e["u"], e["v"], e["w"]], axis=1).astype(np.uint64)
                t = min(cac.intermediate_size(int(ns[i]), len(e)), 4096)
                for k in range(budget_T[i]):
                    oracle.sqrt_trial(int(ns[i]), e3, t, i, k)
            done += 1
            if time.perf_counter() - t0 > cpu_budget_s:
                break
        cpu_s = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": done / cpu_s, "unit": "queries/s", "cores": 1, "kind": "port",
                               "sample": f"the {done} smallest of the {count} queries (n <= {int(np.sort(ns)[done - 1])}) "
                                         f"answered one after another on one core in {cpu_s:.1f} s: CC by the "
                                         f"oracle's union-find, min cut by its square-root trials with the same "
                                         f"budgets; larger queries cost more, so this overstates the CPU rate"}
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only (got --gpus {args.gpus}); "
              "multi-GPU sharding is a design note, not part of this benchmark", file=sys.stderr)
        sys.exit(2)
    sys.path.insert(0, str(ROOT / "oracle"))
    import cac

    scale, m = args.scale, 1 << args.edge_log2
    n = 1 << scale
    th = cac.rmat_thresholds(0.45, 0.22, 0.22)
    ctx = cac.Context(0)
    d_edges = ctx.alloc(m * 8)
    d_lab = ctx.alloc(n * 4)
    t0 = time.perf_counter()
    draws = ctx.gen_rmat_dev(scale, m, args.seed, th, d_edges.ptr)
    gen_s = time.perf_counter() - t0
    print(f"# generated R-MAT scale {scale} m {m} ({draws} draws) in {gen_s:.2f} s", file=sys.stderr, flush=True)

    def measure(step, tag):
        """Warm-up, then exactly args.steps timed steps between two synchronisations."""
        for _ in range(args.warmup):
            step()
        ctx.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            nc = step()
        ctx.synchronize()
        dt = time.perf_counter() - t0
        avg, per_step = time_kernels(ctx, step, args.prof_steps)
        dom = max(per_step, key=per_step.get)
        dom_bytes = kernel_bytes(dom, m, n)
        r = {"ms_per_step": dt * 1e3 / args.steps, "gteps": m / (dt / args.steps) / 1e9, "components": nc,
             "kernels_ms_per_launch": avg, "kernels_ms_per_step": per_step, "dominant": dom,
             "dominant_bytes": dom_bytes, "dominant_avg_ms": avg[dom],
             "dominant_achieved_gbs": dom_bytes / (avg[dom] * 1e-3) / 1e9}
        print(f"# {tag}: {args.steps} steps, {r['ms_per_step']:.3f} ms/step, {nc} components", file=sys.stderr,
              flush=True)
        return r

    # edge-list input: cac_cc_dev, the drop-in for connectedComponents on an edge list
    edge = measure(lambda: ctx.cc_dev(d_edges.ptr, m, n, d_lab.ptr), "edge list")
    lab_edges = d_lab.download(np.uint32, n)
    # symmetric CSR input (built once from the same edge list, outside the timed steps)
    d_off = ctx.alloc((n + 1) * 4)
    d_cols = ctx.alloc(2 * m * 4)
    ctx.synchronize()
    t0 = time.perf_counter()
    nnz = ctx.csr_build_dev(d_edges.ptr, m, n, d_off.ptr, d_cols.ptr)
    csr_build_ms = (time.perf_counter() - t0) * 1e3
    csr = measure(lambda: ctx.cc_csr_dev(d_off.ptr, d_cols.ptr, n, d_lab.ptr), "csr")
    csr["labels_equal_edge_list_path"] = bool(np.array_equal(d_lab.download(np.uint32, n), lab_edges))
    # the same step replayed from a captured HIP graph (cac_cc_csr_graph_dev)
    ctx.cc_csr_graph_dev(d_off.ptr, d_cols.ptr, n, d_lab.ptr)
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        gnc = ctx.cc_csr_graph_dev(d_off.ptr, d_cols.ptr, n, d_lab.ptr)
    ctx.synchronize()
    gdt = (time.perf_counter() - t0) / args.steps
    csr["graph_replay"] = {"ms_per_step": gdt * 1e3, "gteps": m / gdt / 1e9, "components": gnc,
                           "labels_equal_edge_list_path": bool(np.array_equal(d_lab.download(np.uint32, n),
                                                                              lab_edges))}
    csr["csr_build_ms"] = csr_build_ms
    csr["nnz"] = nnz
    del lab_edges
    d_off.free()
    d_cols.free()
    head = csr if args.cc_input == "csr" else edge
    dom = head["dominant"]
    traffic = None
    tfile = ROOT / "profiles" / f"{dom}_rmat{scale}_{args.edge_log2}_traffic.json"
    if tfile.exists():
        try:
            traffic = json.loads(tfile.read_text()).get("hbm_bytes_per_launch")
        except Exception:
            traffic = None
    pipe_bytes = 8.0 * m + 12.0 * n  # the problem's bytes: edge list once + parent init, final read, label write
    pipe_ms = sum(head["kernels_ms_per_step"].values())

    if args.copy_probe_mib > 0:
        copy_probe(ctx, args.copy_probe_mib)

    base = None
    if args.cpu_baseline_seconds > 0:
        e_host = d_edges.download(np.uint32, 2 * m).reshape(m, 2)
        base = cpu_baseline(e_host, n, args.cpu_baseline_seconds)
        del e_host

    held = ("symmetric CSR resident in HBM (Afforest; CSR built from the edge list before timing, "
            f"{csr_build_ms:.0f} ms)" if args.cc_input == "csr" else "edge list resident in HBM")
    out = {
        "metric": METRIC,
        "value": head["gteps"],
        "unit": "GTEPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32",
        "data": "synthetic",
        "config": {
            "workload": f"connected components, R-MAT 2^{scale} vertices / 2^{args.edge_log2} undirected edges, "
                        f"(a,b,c,d)=(0.45,0.22,0.22,0.11), loops+duplicates removed, {held}",
            "vertices": n,
            "edges": m,
            "components": head["components"],
            "label_form": "minimum vertex id",
            "cc_input": args.cc_input,
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dom,
            "achieved": head["dominant_achieved_gbs"],
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": head["dominant_achieved_gbs"] / HBM_PEAK_GBS,
            "traffic": traffic,
            "algorithmic_bytes": head["dominant_bytes"],
            "avg_ms": head["dominant_avg_ms"],
        },
        "kernels_ms": head["kernels_ms_per_launch"],
        "pipeline": {"algorithmic_bytes": pipe_bytes, "kernel_ms": pipe_ms,
                     "achieved_gbs": pipe_bytes / (pipe_ms * 1e-3) / 1e9,
                     "frac": pipe_bytes / (pipe_ms * 1e-3) / 1e9 / HBM_PEAK_GBS},
        "cc_paths": {"edge_list": edge, "csr": csr},
        "cpu_baseline": base,
    }
    d_edges.free()
    d_lab.free()

    wanted = set() if args.extras == "none" else set(args.extras.split(","))
    extras = {}
    for key, fn in (("c1", lambda: bench_c1(ctx, cac, min(args.cpu_baseline_seconds, 10.0))),
                    ("c3", lambda: bench_c3(ctx, cac, args.c3_trials, args.cpu_baseline_seconds)),
                    ("c4", lambda: bench_c4(ctx, cac, args.c4_graphs, args.c4_stride, args.c4_ref_graphs,
                                            min(args.cpu_baseline_seconds, 15.0))),
                    ("c5", lambda: bench_c5(cac, args.c5_queries, args.c5_rate, min(args.cpu_baseline_seconds, 15.0))),
                    ("ks", lambda: bench_ks(ctx, cac, [int(x) for x in args.ks_sizes.split(",")],
                                                    min(args.cpu_baseline_seconds, 20.0))),
                    ("sq", lambda: bench_sq(ctx, cac, [int(x) for x in args.sq_sizes.split(",")],
                                                    min(args.cpu_baseline_seconds, 15.0))),
                    ("rm", lambda: bench_rm(ctx, cac, [int(x) for x in args.rm_degrees.split(",")])),
                    ("ap", lambda: bench_ap(ctx, cac, [int(x) for x in args.sq_sizes.split(",")],
                                                    min(args.cpu_baseline_seconds, 10.0)))):
        if key in wanted:
            # every leg starts on a fresh context: a context keeps the buffers
            # it grew (the dense Karger-Stein levels of c3 hold ~160 GB), and
            # the served leg's workers need that memory
            ctx.close()
            ctx = cac.Context(0)
            t0 = time.perf_counter()
            extras[key] = fn()
            print(f"# {key} done in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    if extras:
        out["extras"] = extras
    ctx.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
