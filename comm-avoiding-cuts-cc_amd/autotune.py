"""Autotuner for the launch shapes of libcac.so on the device it runs on.

Knob CAC_TUNE_KS_CONTRACT_BLOCK: threads per workgroup of the dense
Karger-Stein contraction (k_ks_contract<B>), per row-size class (key 512 holds
children of at most 512 vertices, keys 1024, 2048 and 4096 hold s in
(key/2, key]).  The built-in rule takes the smallest power of two with at most
eight row entries per thread; larger blocks give each thread fewer entries and
a workgroup more waves, which pays or not depending on how many children of
that size run at once and on their row length.

For each class, largest first, every candidate block size (powers of two from
the class minimum to 1024) is timed on a dense Karger-Stein solve whose top
levels lie in that class: a complete graph with n = key vertices and weights
U[1, 42], the classes above it held at their tuned values and the ones below at
the built-in rule.  The fastest candidate whose raw per-trial cuts equal the
built-in rule's wins (results do not depend on the block size; a candidate that
changed one would be a bug, and the tuner stops).

Knob CAC_TUNE_SAMPLE_OCCUPANCY: 1024-thread sampling workgroups per CU (2:
64 registers per thread with some spilled, 1: the registers the kernel wants)
for a launch of more square-root trials than CUs, per edge-count class (keys
16, 20, 24: m < 2^16, 2^20, 2^24).  Each class is timed on many square-root
trials of one graph of that size (the trials of a batch or of one large solve
run that way), and the faster setting whose raw cuts equal the built-in
setting's wins.

Knob CAC_TUNE_KS_COMPACT_BLOCK: threads per workgroup (256, 512, 1024) of the
dense Karger-Stein compaction, per child-size class (keys 512, 1024, 2048,
4096, 8192: t' <= 512, then t' in (key/2, key]).  One solve of 57 trials on the
complete graph n = 8192 holds children of every class; each class is swept,
largest first, with the classes above at their tuned values and the ones below
at the built-in 512, and the fastest setting whose raw cuts equal the built-in
rule's wins.

Writes {"device": ..., "ks_contract_block": {"512": B, ...},
"sample_occupancy": {"16": k, ...}, "ks_compact_block": {"512": B, ...},
"timings_ms": ...} to tuning/gfx950.json
next to this file (or --out); every cac.Context applies that file when it
exists.  Run on the target GPU:

    python comm-avoiding-cuts-cc_amd/autotune.py [--out PATH] [--reps 3]
"""
from __future__ import annotations

import argparse
import json
import sys
import time
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent))
import cac  # noqa: E402

# (class key, trials per timed solve).  Enough trials that thousands of children
# of each class run at once, as at the deep levels of a large solve (169 trials
# at n = 8192 run 2^10 times more children of 260 vertices than the root has):
# with only hundreds, larger blocks look as fast, though they cut occupancy
# where it matters.  About 0.2-1 s per solve on one MI355X.
CLASSES = ((4096, 128), (2048, 512), (1024, 2048), (512, 4096))
BLOCKS = (64, 128, 256, 512, 1024)


def candidates(key: int):
    lo = max(64, key // 8)
    return [b for b in BLOCKS if b >= lo]


def time_solve(ctx, d_adj: int, n: int, trials: int, reps: int):
    raw, _ = ctx.ks_dense_dev(d_adj, n, 0, 0, trials)  # warm-up (grows the level buffers)
    times = []
    for _ in range(reps):
        ctx.synchronize()
        t0 = time.perf_counter()
        r, _ = ctx.ks_dense_dev(d_adj, n, 0, 0, trials)
        times.append((time.perf_counter() - t0) * 1e3)
        if r.tolist() != raw.tolist():
            raise RuntimeError("repeated solve changed a raw cut")
    times.sort()
    return times[len(times) // 2], raw


def tune(device: int = 0, reps: int = 3, log=print):
    table, timings = {}, {}
    with cac.Context(device, tuning=False) as ctx:
        for key, trials in CLASSES:
            n = key
            d_adj = ctx.alloc(n * n * 8)
            ctx.gen_complete_dev(n, 0, 42, d_adj.ptr)
            for k2, b2 in table.items():  # classes above: their tuned values
                ctx.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, int(k2), b2)
            ctx.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, key, 0)
            base_ms, base_raw = time_solve(ctx, d_adj.ptr, n, trials, reps)
            res = {}
            for b in candidates(key):
                ctx.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, key, b)
                ms, raw = time_solve(ctx, d_adj.ptr, n, trials, reps)
                if raw.tolist() != base_raw.tolist():
                    raise RuntimeError(f"block {b} for class {key} changed a raw cut")
                res[b] = ms
                log(f"class {key}: block {b:4d}  {ms:8.2f} ms  ({trials} trials, n={n})")
            best = min(res, key=res.get)
            table[str(key)] = best
            timings[str(key)] = {"n": n, "trials": trials, "built_in_ms": base_ms,
                                 "ms": {str(b): v for b, v in res.items()}}
            log(f"class {key}: picked {best} (built-in rule {base_ms:.2f} ms, picked {res[best]:.2f} ms)")
            ctx.tune_set(cac.TUNE_KS_CONTRACT_BLOCK, key, 0)
            d_adj.free()
    return table, timings


# (class key, graph, trials): one graph per edge-count class and enough trials
# to fill the device several times over.
SAMPLE_CLASSES = ((16, ("er", 2000, 20.0), 4096), (20, ("ws", 32000, 16), 2048), (24, ("er", 100000, 100.0), 512))


def _sample_graph(spec):
    kind, n, d = spec
    if kind == "ws":
        return n, cac.gen_ws(n, int(d), 0)
    return n, cac.gen_gnp(n, d, 0, 0, wmax=100)


def tune_sampling(device: int = 0, reps: int = 3, log=print):
    table, timings = {}, {}
    with cac.Context(device, tuning=False) as ctx:
        for key, spec, trials in SAMPLE_CLASSES:
            n, e = _sample_graph(spec)
            t = cac.intermediate_size(n, len(e))
            res, raw0 = {}, None
            for occ in (2, 1):
                ctx.tune_set(cac.TUNE_SAMPLE_OCCUPANCY, key, occ)
                raw = ctx.sqrt_cut_trials(e, n, t, 0, 0, trials)  # warm-up
                times = []
                for _ in range(reps):
                    t0 = time.perf_counter()
                    r = ctx.sqrt_cut_trials(e, n, t, 0, 0, trials)
                    times.append((time.perf_counter() - t0) * 1e3)
                    if r.tolist() != raw.tolist():
                        raise RuntimeError("repeated trials changed a raw cut")
                if raw0 is None:
                    raw0 = raw
                elif raw.tolist() != raw0.tolist():
                    raise RuntimeError(f"occupancy {occ} for class {key} changed a raw cut")
                res[occ] = sorted(times)[len(times) // 2]
                log(f"sampling class 2^{key}: {occ} workgroup(s) per CU  {res[occ]:8.2f} ms  "
                    f"({trials} trials, n={n}, m={len(e)}, t={t})")
            best = min(res, key=res.get)
            table[str(key)] = best
            timings[str(key)] = {"n": n, "m": len(e), "t": t, "trials": trials,
                                 "ms": {str(k): v for k, v in res.items()}}
            ctx.tune_set(cac.TUNE_SAMPLE_OCCUPANCY, key, 0)
    return table, timings


COMPACT_BLOCKS = (256, 512, 1024)


def tune_compact(device: int = 0, reps: int = 3, log=print):
    n, trials = 8192, 57
    table, timings = {}, {}
    with cac.Context(device, tuning=False) as ctx:
        d_adj = ctx.alloc(n * n * 8)
        ctx.gen_complete_dev(n, 0, 42, d_adj.ptr)
        base_ms, base_raw = time_solve(ctx, d_adj.ptr, n, trials, reps)
        for key in reversed(cac.TUNE_KS_COMPACT_KEYS):
            res = {}
            for b in COMPACT_BLOCKS:
                ctx.tune_set(cac.TUNE_KS_COMPACT_BLOCK, key, b)
                ms, raw = time_solve(ctx, d_adj.ptr, n, trials, reps)
                if raw.tolist() != base_raw.tolist():
                    raise RuntimeError(f"compaction block {b} for class {key} changed a raw cut")
                res[b] = ms
                log(f"compaction class {key}: block {b:4d}  {ms:8.2f} ms  ({trials} trials, n={n})")
            best = min(res, key=res.get)
            table[str(key)] = best
            ctx.tune_set(cac.TUNE_KS_COMPACT_BLOCK, key, best)
            timings[str(key)] = {"n": n, "trials": trials, "built_in_ms": base_ms,
                                 "ms": {str(b): v for b, v in res.items()}}
            log(f"compaction class {key}: picked {best}")
        d_adj.free()
    return table, timings


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--out", default=str(cac.TUNING_PATH))
    ap.add_argument("--device", type=int, default=0)
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--only", choices=("all", "sampling", "compact"), default="all",
                    help="tune only that knob and keep the committed tables of the others")
    args = ap.parse_args()
    ap_log = lambda s: print(s, flush=True)  # noqa: E731
    old = json.loads(Path(args.out).read_text()) if args.only != "all" and Path(args.out).exists() else {}
    old_t = old.get("timings_ms", {})
    if args.only in ("all",) or "ks_contract_block" not in old:
        table, timings = tune(args.device, args.reps, log=ap_log)
    else:
        table, timings = old["ks_contract_block"], old_t.get("ks_contract_block", {})
    if args.only in ("all", "sampling") or "sample_occupancy" not in old:
        stable, stimings = tune_sampling(args.device, args.reps, log=ap_log)
    else:
        stable, stimings = old["sample_occupancy"], old_t.get("sample_occupancy", {})
    if args.only in ("all", "compact") or "ks_compact_block" not in old:
        ctable, ctimings = tune_compact(args.device, args.reps, log=ap_log)
    else:
        ctable, ctimings = old["ks_compact_block"], old_t.get("ks_compact_block", {})
    out = {"device": "gfx950 (MI355X)", "knobs": ["ks_contract_block", "sample_occupancy", "ks_compact_block"],
           "ks_contract_block": table, "sample_occupancy": stable, "ks_compact_block": ctable,
           "timings_ms": {"ks_contract_block": timings, "sample_occupancy": stimings, "ks_compact_block": ctimings}}
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text(json.dumps(out, indent=1) + "\n")
    print(f"wrote {args.out}: {table}", flush=True)


if __name__ == "__main__":
    main()
