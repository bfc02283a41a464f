#!/usr/bin/env python3
"""Benchmark of the MI355X DP backend.

Workload (one "step"): batched Zuker MFE + backtrace of 4096 RNA sequences,
lengths log-uniform in [50, 2000], i.i.d. 50% GC (seed 3), one persistent
tile-queue launch plus one finalize launch.  Prints one JSON line.

Served stream (reported beside it under "served"): Poisson arrivals at a fixed
rate of Zuker fold requests (lengths log-uniform 30-3000, 50% GC) and Gotoh SW
requests (each sequence log-uniform 500-20000 bp, ACGT, +2/-3/-5/-2, traced
from the best cell) in a 4:1 mix, plus matrix-chain requests (64-512
matrices, dimensions uniform in [1,1000]) as 10% of all requests, through the
continuous-batching server under a fixed HBM budget: p50/p99 latency, GCUPS,
admission rejects, and a bit-exact check of a sample against the same
requests solved alone.

value = cell updates per second of the batched step, with inputs resident in
HBM: the device time of the fill and finalize kernels (HIP events on the
library's stream).  Also reported: the same sequences as plain per-instance
launches (a fixed 1-in-16 sample), the end-to-end rate of the C call
including PCIe copies, the roofline of the fill kernel and a single-thread
CPU baseline (the oracle, a bounded sample of the same sequences).

--workload selects one of the single-instance configurations instead
(zuker2048, rnafold2048, nussinov1024, swat16384); each prints its own JSON
line.
"""
import argparse
import json
import math
import os
import sys
import threading
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lamp-dp-mt_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "GCUPS batched/served vs plain per-instance (bit-exact); p50/p99 fold latency"
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz


def cells(n):
    return (n + 1) * (n + 2) // 2


def _tet(m):
    return m * (m + 1) * (m + 2) // 6 if m > 0 else 0


def x_candidates(n):
    """Split points k in [i+7, j-7] of one ml/ml1 min-plus term over all cells:
    sum over diagonals d=j-i of (n-d+1)*(d-13)."""
    return _tet(n - 13)


def ml_candidates(n):
    """Split candidates of the reference's ml/ml1 k-loops over all cells
    (zu_kern.h:92-184): rules 7 and 9, k in [i+7, j-7], and rule 11,
    k in [i+7, j-1]."""
    return 2 * _tet(n - 13) + _tet(n - 7)


def load_profile_value(name, key):
    """A `key: value` line of a committed profile summary (profiles/<name>)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith(key + ":"):
                return float(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def load_traffic(name):
    """HBM bytes per launch measured with rocprofv3 FETCH_SIZE / WRITE_SIZE
    (scaled by the copy-kernel calibration), from the committed summary."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch:"):
                return int(line.split(":")[1]), "profiles/" + name
    except OSError:
        pass
    return None, None


def pct(xs, p):
    xs = sorted(xs)
    if not xs:
        return None
    k = min(len(xs) - 1, max(0, int(math.ceil(p / 100.0 * len(xs))) - 1))
    return xs[k]


def served_requests(rate, seconds, seed=5):
    """Seeded Poisson stream of the served mix: (arrival s, kind, payload)."""
    from dpx import gen
    return gen.served_stream(rate, seconds, seed)


def golden_sample(seqs, results, traces):
    """Compares the sampled instances of the 4096-sequence batch with the
    committed answers of the reference's bench2/zu.c (score and trace; the
    second position of a rule with one split position is not compared).
    Returns (all equal, instances compared); (None, 0) for another batch."""
    path = os.path.join(ROOT, "tests", "golden", "zuker_batch4096_sample.json")
    with open(path) as f:
        fx = json.load(f)
    if len(seqs) != 4096 or any(seqs[e["index"]] != e["seq"] for e in fx):
        return None, 0
    tl = [2] * 6 + [1] * 9
    norm = lambda tr: [[a, b, r, p0, p1 if tl[r] > 1 else 0] for a, b, r, p0, p1 in tr]
    ok = all(results[e["index"]] == e["score"] and norm(traces[e["index"]]) == e["trace"] for e in fx)
    return ok, len(fx)


SW_SCORES = (2, -3, -5, -2)


def serve(ctx, rate, seconds, budget, slots, max_batch, check_every):
    import dpx
    reqs = served_requests(rate, seconds)
    srv = dpx.Server(ctx, hbm_budget=budget, slots=slots, max_batch=max_batch)
    # load every kernel once before the clock starts
    w = [srv.submit_zuker("GGGGAAAACCCC"), srv.submit_swat("ACGT" * 20, "AGCT" * 20, *SW_SCORES, trace_from=1),
         srv.submit_matchain([3, 4, 5], [4, 5, 6])]
    srv.drain(w)
    ids = [None] * len(reqs)
    rejected = []
    check = set(range(0, len(reqs), check_every)) if check_every > 0 else set()
    keep = set()  # ids whose traces are fetched (the bit-exact sample); the others are dropped unread
    known = set()

    def produce():
        t0 = dpx.now()
        for q, (at, kind, x) in enumerate(reqs):
            d = t0 + at - dpx.now()
            if d > 0:
                time.sleep(d)
            try:
                if kind == "zuker":
                    ids[q] = srv.submit_zuker(x)
                elif kind == "swat":
                    ids[q] = srv.submit_swat(x[0], x[1], *SW_SCORES, trace_from=1)
                else:
                    ids[q] = srv.submit_matchain(x[0], x[1])
            except dpx.DpxError as e:
                rejected.append((q, e.code))
                continue
            if q in check:
                keep.add(ids[q])
            known.add(ids[q])  # after keep: an id in known has its keep decision made

    th = threading.Thread(target=produce)
    t_start = dpx.now()
    th.start()
    got = {}
    undecided = []  # finished before the producer recorded its id

    def take(res):
        tr = None
        if res.id in keep:
            tr = srv.trace(res.id, res.trace_len)
        else:
            srv.drop_trace(res.id)
        got[res.id] = (res, tr)

    while th.is_alive() or len(got) + len(undecided) < sum(1 for x in ids if x is not None):
        for res in srv.poll(wait_us=5000):
            if res.trace_len <= 0:
                got[res.id] = (res, None)
            elif res.id in known:
                take(res)
            else:
                undecided.append(res)
    th.join()
    for res in undecided:
        take(res)
    st = srv.stats()
    srv.close()
    done = [got[i][0] for i in ids if i is not None]
    t_end = max(r.t_done for r in done)
    cells = sum(r.cells for r in done)
    lat = {}
    for k in ("zuker", "swat", "matchain"):
        lat[k] = [(got[ids[q]][0].t_done - got[ids[q]][0].t_submit) * 1e3
                  for q in range(len(reqs)) if reqs[q][1] == k and ids[q] is not None]
    every = [x for v in lat.values() for x in v]
    # bit-exact sample: the same requests solved alone
    exact, checked = True, 0
    for q in sorted(check):
        if ids[q] is None:
            continue
        res, tr = got[ids[q]]
        kind, x = reqs[q][1], reqs[q][2]
        if kind == "zuker":
            want = ctx.zuker(x)
            ok = (res.score, tr) == want
        elif kind == "swat":
            s, best, wtr = ctx.swat(x[0], x[1], *SW_SCORES, trace_from=1)
            ok = (res.score, (res.best, res.best_i, res.best_j), tr) == (s, best, wtr)
        else:
            ok = (res.score, tr) == ctx.matchain(x[0], x[1])
        exact &= bool(ok) and res.status == 0
        checked += 1
    failed = sum(1 for r in done if r.status != 0)
    return {
        "gcups": round(cells / (t_end - t_start) / 1e9, 4),
        "arrival_rate_per_s": rate, "seconds": seconds, "requests": len(reqs),
        "completed": len(done) - failed, "failed": failed, "rejected": len(rejected),
        "latency_ms": {k: {"p50": round(pct(v, 50), 3), "p99": round(pct(v, 99), 3), "n": len(v)}
                       for k, v in lat.items() if v},
        "latency_ms_all": {"p50": round(pct(every, 50), 3), "p99": round(pct(every, 99), 3)},
        "cells": cells, "span_s": round(t_end - t_start, 3),
        "hbm_budget_bytes": budget, "slots": slots, "max_batch": max_batch,
        "launches": st["launches"], "mean_batch": round(len(done) / max(1, st["launches"] - 3), 2),
        "peak_bytes_in_flight": st["peak_bytes_in_flight"],
        "bitexact_sample": exact, "checked": checked,
        "mix": "Zuker:SW 4:1 (72%:18%), matrix chain 10% of requests",
    }


def serve_sweep(ctx, args, budget):
    """Offered load swept upwards: the achieved GCUPS levels off at the
    server's capacity (saturation); then p50/p99 at 80% of it (the arrival
    rate whose offered cells are 80% of the saturation GCUPS)."""
    from dpx import gen
    rates = [float(r) for r in args.serve_sweep.split(",")]
    pts = []
    for rate in rates:
        r = serve(ctx, rate, args.serve_sweep_seconds, budget, args.serve_slots, 256, 0)
        pts.append({"rate": rate, "gcups": r["gcups"], "completed": r["completed"], "requests": r["requests"],
                    "span_s": r["span_s"], "latency_ms_all": r["latency_ms_all"], "mean_batch": r["mean_batch"]})
        print("bench: sweep rate %.0f/s: %.3f GCUPS, p50 %.1f ms, p99 %.1f ms, span %.2f s" % (
            rate, r["gcups"], r["latency_ms_all"]["p50"], r["latency_ms_all"]["p99"], r["span_s"]),
            file=sys.stderr, flush=True)
    sat = max(p["gcups"] for p in pts)
    # mean cells per request of the stream (same generator, long sample)
    sample = gen.served_stream(1000.0, 20.0, 9)
    mean_cells = sum((len(x) + 1) * (len(x) + 2) // 2 if k == "zuker" else
                     (len(x[0]) + 1) * (len(x[1]) + 1) if k == "swat" else (len(x[0]) + 1) * (len(x[0]) + 2) // 2
                     for _, k, x in sample) / len(sample)
    rate80 = 0.8 * sat * 1e9 / mean_cells
    at80 = serve(ctx, rate80, args.serve_sweep_seconds * 2, budget, args.serve_slots, 256, args.serve_check_every)
    print("bench: at 80%% of saturation (%.0f/s): %.3f GCUPS, p50 %.1f ms, p99 %.1f ms" % (
        rate80, at80["gcups"], at80["latency_ms_all"]["p50"], at80["latency_ms_all"]["p99"]), file=sys.stderr, flush=True)
    return {"points": pts, "saturation_gcups": sat, "mean_cells_per_request": round(mean_cells),
            "rate_at_80pct": round(rate80, 1),
            "at_80pct": {k: at80[k] for k in ("gcups", "requests", "completed", "failed", "rejected", "latency_ms",
                                               "latency_ms_all", "mean_batch", "bitexact_sample", "checked", "span_s")}}


def single(args):
    """One instance of a single-GPU configuration, timed over --steps runs."""
    import dpx
    from dpx import gen
    import oracle_lib as o
    ctx = dpx.Context()
    line = {"n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "int32", "data": "synthetic", "unit": "GCUPS"}
    if args.workload.startswith("grammar-"):
        return single_grammar(args, ctx, line)
    if args.workload == "swat16384":
        a, b = gen.dna_pair(16384, 16384, 1)
        run = lambda: ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
        n_cells = 16385 * 16385
        line["config"] = {"workload": "Smith-Waterman/Gotoh 16384 x 16384, i.i.d. ACGT seed 1, +2/-3/-5/-2, "
                                      "best local score + traceback from the best cell"}
        if args.swat_trace_budget is not None:
            # below the code size (128 MiB here) the traceback recomputes segments from checkpoint rows
            ctx.set_swat_trace_budget(args.swat_trace_budget)
            line["config"]["swat_trace_budget_bytes"] = args.swat_trace_budget
        check = lambda r: (r[0], r[1], [t[:3] for t in r[2]]) == (lambda x: (x[0], x[1], [t[:3] for t in x[2]]))(
            o.swat(a, b, 2, -3, -5, -2, from_best=True))
    elif args.workload == "nussinov1024":
        s = gen.rna_uniform(1024, 0)
        run = lambda: ctx.nussinov(s, min_loop=3)
        n_cells = cells(1024)
        line["config"] = {"workload": "Nussinov n=1024, i.i.d. uniform ACGU seed 0, min hairpin loop 3, score + backtrace"}
        check = lambda r: (r[0], [t[:4] for t in r[1]]) == (lambda x: (x[0], [t[:4] for t in x[1]]))(o.nussinov(s, min_loop=3))
    elif args.workload == "rnafold2048":
        s = gen.rna_gc(2048, 2)
        run = lambda: ctx.rnafold(s)
        n_cells = cells(2048)
        line["config"] = {"workload": "RNAfold grammar MFE n=2048, i.i.d. 50% GC seed 2, Turner 2004, MFE + backtrace"}
        check = None  # the oracle takes about 20 s here; the GPU tests pin this input against a committed fixture
    else:
        s = gen.rna_gc(2048, 2)
        run = lambda: ctx.zuker(s)
        n_cells = cells(2048)
        line["config"] = {"workload": "Zuker MFE n=2048, i.i.d. 50% GC seed 2, Turner 2004, MFE + backtrace"}
        check = None  # O(n^3) oracle run is minutes; the GPU tests pin n=2048 against the reference fixture
    for _ in range(args.warmup):
        r = run()
    fills, fins = [], []
    for _ in range(args.steps):
        r = run()
        f, z = ctx.last_timing()
        fills.append(f)
        fins.append(z)
        print("bench: step fill=%.3fms finalize=%.3fms" % (f, z), file=sys.stderr, flush=True)
    ms = (sum(fills) + sum(fins)) / args.steps
    line.update({"metric": "GCUPS (%s)" % args.workload, "value": round(n_cells / (ms * 1e-3) / 1e9, 4),
                 "ms_per_step": round(ms, 4), "fill_ms": round(sum(fills) / args.steps, 4),
                 "finalize_ms": round(sum(fins) / args.steps, 4), "cells": n_cells,
                 "bitexact": check(r) if (check and not args.no_cpu) else None})
    if args.workload == "swat16384":
        fill_s = sum(fills) / args.steps * 1e-3
        traffic, tsrc = load_traffic("sw_fill_swat_16384x16384.txt")
        valu_util, _ = load_profile_value("sw_fill_swat_16384x16384.txt", "valu_issue_util")
        # the band pipeline's ceilings: its critical path is n1 + n2 + nbands * chunk
        # steps; a step needs at least the dependent chain of its cell (wave
        # shift, add, clamp, max with the extension, with the pair, with g2:
        # about 30 cycles) and, with one wave per band, the issue of all its
        # instructions (VALU per step from the counters x 4 cycles)
        chunk = ctx.get_tuning("swat", 7)["chunk"] or 32
        steps = 16384 + 16384 + (16384 // 64) * chunk
        dep_ms = steps * 30 / 2.4e9 * 1e3
        issue_ms = steps * 39.3 * 4 / 2.4e9 * 1e3
        line["roofline"] = {"bound": "valu", "achieved": round(n_cells * 10 / fill_s / 1e12, 4),
                            "peak": round(PEAK_INT32_TOPS, 2), "unit": "TOP/s",
                            "frac": round(n_cells * 10 / fill_s / 1e12 / PEAK_INT32_TOPS, 5),
                            "ops": "about 10 int32 ops per cell", "traffic": traffic, "traffic_source": tsrc,
                            "hbm_frac": round(n_cells * 0.5 / fill_s / 8e12, 5),
                            "valu_issue_util_counter": valu_util,
                            "critical_path_steps": steps,
                            "dependency_ceiling_ms": round(dep_ms, 3),
                            "one_wave_issue_ceiling_ms": round(issue_ms, 3),
                            "frac_of_issue_ceiling": round(issue_ms / (fill_s * 1e3), 3)}
    print(json.dumps(line))


def single_grammar(args, ctx, line):
    """A reference example grammar through the generic path (compiled to HIP
    at run time), with the built-in kernel of the same grammar beside it."""
    import random
    import dpx.examples as G
    from dpx import gen
    w = args.workload
    if w == "grammar-nussinov1024":
        s = gen.rna_uniform(1024, 0)
        g = G.nussinov()
        run_in = (s,)
        builtin = lambda: ctx.nussinov(s, min_loop=0)
        n_cells = cells(1024)
        line["config"] = {"workload": "Nussinov.scala grammar, n=1024, i.i.d. uniform ACGU seed 0, score + backtrace, "
                                      "generic path (dpx_grammar_parse)"}
    elif w == "grammar-zuker2048":
        s = gen.rna_gc(2048, 2)
        g = G.zuker()
        run_in = (s,)
        builtin = lambda: ctx.zuker(s)
        n_cells = cells(2048)
        line["config"] = {"workload": "Zuker.scala grammar (ZukerMFE, librna on the device), n=2048, i.i.d. 50% GC seed 2, "
                                      "Turner 2004, MFE + backtrace, generic path"}
    elif w == "grammar-nussinov-batch":
        seqs = gen.rna_batch(1024, 50, 500, 5, gc=0.5)
        g = G.nussinov()
        run_in = None
        builtin = lambda: ctx.nussinov_batch(seqs, min_loop=0)
        n_cells = sum(cells(len(x)) for x in seqs)
        line["config"] = {"workload": "Nussinov.scala grammar, batch of 1024 sequences, lengths log-uniform [50,500] "
                                      "(seed 5), scores + backtraces, generic path (dpx_grammar_parse_batch)"}
    elif w == "grammar-chain1024":
        rng = random.Random(1)
        dims = [rng.randint(1, 100) for _ in range(1025)]
        mats = [(dims[q], dims[q + 1]) for q in range(1024)]
        g = G.matrix_chain()
        run_in = (mats,)
        builtin = lambda: ctx.matchain(dims[:-1], dims[1:])
        n_cells = cells(1024)
        line["config"] = {"workload": "MatrixMult2.scala grammar, 1024 matrices (dims uniform 1..100, seed 1), "
                                      "cost + backtrace, generic path"}
    else:  # grammar-swat4000
        a, b = gen.dna_pair(4000, 4000, 1)
        g = G.swat_affine(2, -3, -5, -2)
        run_in = (a, b)
        builtin = lambda: ctx.swat(a, b, 2, -3, -5, -2)
        n_cells = 4001 * 4001
        line["config"] = {"workload": "SWatAffine.scala grammar, 4000 x 4000 i.i.d. ACGT seed 1, +2/-3/-5/-2, "
                                      "M(n1,n2) + backtrace, generic path"}
    t0 = time.perf_counter()
    cg = ctx.grammar(g)
    compile_s = time.perf_counter() - t0
    run = (lambda: cg.parse_batch(seqs)) if run_in is None else (lambda: cg.parse(*run_in))
    for _ in range(args.warmup):
        r = run()
    fills = []
    for _ in range(args.steps):
        r = run()
        fills.append(ctx.last_timing()[0])
        print("bench: step fill=%.3fms" % fills[-1], file=sys.stderr, flush=True)
    builtin()
    bres = builtin()
    bfill = sum(ctx.last_timing())
    ms = sum(fills) / len(fills)
    if run_in is None:
        same = [x[0] for x in r] == list(bres[0]) and [x[2] for x in r] == list(bres[1])
    else:
        same = (r[0] == bres[0]) and [t[:3] for t in r[2]] == [t[:3] for t in bres[-1]]
    line.update({"metric": "GCUPS (%s)" % w, "value": round(n_cells / (ms * 1e-3) / 1e9, 4), "ms_per_step": round(ms, 4),
                 "fill_ms": round(ms, 4), "cells": n_cells, "compile_s": round(compile_s, 3),
                 "builtin_ms": round(bfill, 4), "same_as_builtin": bool(same)})
    print(json.dumps(line))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="batch", choices=["batch", "zuker2048", "rnafold2048", "nussinov1024", "swat16384",
                                                            "grammar-nussinov1024", "grammar-chain1024", "grammar-swat4000",
                                                            "grammar-nussinov-batch", "grammar-zuker2048"])
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--plain-stride", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--autotune", action="store_true", help="re-run the autotuner instead of loading its table")
    ap.add_argument("--swat-trace-budget", type=int, default=None,
                    help="swat16384: bytes of traceback codes kept before checkpointed recompute")
    ap.add_argument("--serve-rate", type=float, default=300.0)
    ap.add_argument("--serve-seconds", type=float, default=10.0)
    ap.add_argument("--serve-budget-gib", type=float, default=16.0)
    ap.add_argument("--serve-slots", type=int, default=6)
    ap.add_argument("--serve-sweep", default="1000,4000,8000,12000,16000",
                    help="arrival rates (requests/s) of the saturation sweep; empty to skip")
    ap.add_argument("--serve-sweep-seconds", type=float, default=3.0)
    ap.add_argument("--serve-check-every", type=int, default=25)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: runs on one GPU; --gpus %d is not supported "
              "(a batch's instances would shard across GPUs without communication)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import dpx
    from dpx import gen

    if args.workload != "batch":
        single(args)
        return

    seqs = gen.rna_batch(args.count, 50, 2000, 3, gc=0.5)
    total_cells = sum(cells(len(s)) for s in seqs)
    total_x = sum(x_candidates(len(s)) for s in seqs)
    print("bench: %d sequences, %.3e cells" % (len(seqs), total_cells), file=sys.stderr, flush=True)

    # launch shape per length bucket: the context loads the table the
    # autotuner produced on MI355X; --autotune re-runs the autotuner here
    ctx = dpx.Context()
    if args.autotune:
        t0 = time.perf_counter()
        tune_rep = ctx.autotune("zuker", reps=1)
        tuning = {"source": "dpx_autotune in this run (%.2f s)" % (time.perf_counter() - t0),
                  "zuker": {"b%d(%d-%d)" % (r["bucket"], r["n_lo"], r["n_hi"]): r["wg_per_cu"]
                            for r in tune_rep if r["chosen"]}}
    else:
        tuning = {"source": os.path.relpath(dpx.TUNING_MI355X, ROOT) + " (dpx_autotune on MI355X)",
                  "zuker": {"b%d" % b: ctx.get_tuning("zuker", b)["wg_per_cu"] for b in range(dpx.TUNE_BUCKETS)}}
    print("bench: tuning %s" % json.dumps(tuning), file=sys.stderr, flush=True)
    # eager launches: warm-up, the fill / finalize split and the C call time
    fills, fins, walls = [], [], []
    results = traces = None
    for _ in range(max(1, args.warmup)):
        results, traces = ctx.zuker_batch(seqs)
        f, z = ctx.last_timing()
        fills.append(f)
        fins.append(z)
        walls.append(ctx.last_call_s)
        print("bench: eager fill=%.2fms finalize=%.2fms C call=%.2fms" % (f, z, walls[-1] * 1e3),
              file=sys.stderr, flush=True)
    # timed steps: the batch captured once as a HIP graph (queue reset +
    # persistent fill + finalize), inputs resident in HBM, replayed K times
    plan = ctx.plan(seqs)
    plan.run()
    steps, g_fill, g_fin = [], [], []
    for _ in range(args.steps):
        steps.append(plan.run())
        f, z = plan.timing()  # events before, between and after the fill and finalize graphs
        g_fill.append(f)
        g_fin.append(z)
        print("bench: step (graph replay) %.2fms: fill %.2fms finalize %.2fms" % (steps[-1], f, z),
              file=sys.stderr, flush=True)
    graph_same = plan.results() == (results, traces)
    plan.close()
    dev_ms = sum(steps) / args.steps
    value = total_cells / (dev_ms * 1e-3) / 1e9
    wall_ms = sum(walls) / len(walls) * 1e3
    # the fill kernel's duration over the timed replays
    fill_ms = sum(g_fill) / len(g_fill)

    # plain per-instance launches over a fixed sample of the same sequences
    plain = None
    if args.plain_stride > 0:
        sample = seqs[::args.plain_stride]
        p_cells, p_ms = 0, 0.0
        for s in sample:
            r, _ = ctx.zuker(s)
            f, z = ctx.last_timing()
            p_cells += cells(len(s))
            p_ms += f + z
        plain = p_cells / (p_ms * 1e-3) / 1e9
        print("bench: plain per-instance %.3f GCUPS over %d sequences" % (plain, len(sample)), file=sys.stderr, flush=True)

    # bit-exact check of a stratified sample (every length bucket and the 8
    # longest sequences) against the reference's own generated CPU code's
    # answers for the same sequences, committed as fixtures
    exact, n_exact = golden_sample(seqs, results, traces)

    # algorithmic work of the reference's formulation: 3 int32 ops per
    # ml/ml1 split candidate (rules 7, 9, 11: add, compare, select of the
    # value and its back-pointer, as the reference's generated loop does);
    # the loop-energy candidates of cl are not counted, so this is a lower
    # bound of the work per cell
    cand = sum(ml_candidates(len(s)) for s in seqs)
    ops = 3.0 * cand
    achieved = ops / (fill_ms * 1e-3) / 1e12
    traffic, tsrc = load_traffic("tri_fill_zuker_batch4096_hbm.txt")
    valu_util, vsrc = load_profile_value("tri_fill_zuker_batch4096_counters.txt", "valu_issue_util")
    roofline = {"bound": "valu", "achieved": round(achieved, 4), "peak": round(PEAK_INT32_TOPS, 2), "unit": "TOP/s",
                "frac": round(achieved / PEAK_INT32_TOPS, 5), "traffic": traffic, "traffic_source": tsrc,
                "kernel": "tri_fill<GZu>",
                "ops": "3 int32 ops (add, compare, select) per split candidate of the reference's ml/ml1 loops "
                       "(rules 7, 9, 11): %.4e candidates, %.4e ops per launch" % (cand, ops),
                "frac_2op": round(2.0 * cand / (fill_ms * 1e-3) / 1e12 / PEAK_INT32_TOPS, 5),
                "executed_ops_frac": round(2.0 * total_x / (fill_ms * 1e-3) / 1e12 / PEAK_INT32_TOPS, 5),
                "valu_issue_util_counter": valu_util, "counter_source": vsrc}

    served = None
    if not args.no_serve:
        budget = int(args.serve_budget_gib * (1 << 30))
        served = serve(ctx, args.serve_rate, args.serve_seconds, budget, args.serve_slots, 256, args.serve_check_every)
        print("bench: served %s" % json.dumps(served), file=sys.stderr, flush=True)
        if args.serve_sweep:
            served["sweep"] = serve_sweep(ctx, args, budget)

    cpu = None
    if not args.no_cpu:
        import oracle_check  # the timed CPU baseline (the oracle)
        cpu = oracle_check.cpu_baseline(seqs, args.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dev_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batched Zuker MFE + backtrace: %d RNA sequences, lengths log-uniform [50,2000], "
                               "50%% GC, seed 3, Turner 2004 parameters" % len(seqs),
                   "instances": len(seqs), "cells": total_cells,
                   "parallelism": "single GPU, one persistent tile-queue launch captured as a HIP graph"},
        "plain_per_instance_gcups": round(plain, 4) if plain else None,
        "speedup_vs_plain": round(value / plain, 2) if plain else None,
        "c_call_ms_per_step_pcie": round(wall_ms, 3), "fill_ms": round(fill_ms, 3),
        "finalize_ms": round(sum(g_fin) / len(g_fin), 3),
        "eager_fill_ms": round(sum(fills) / len(fills), 3), "eager_finalize_ms": round(sum(fins) / len(fins), 3),
        "tuning": tuning, "bitexact_sample": exact,
        "bitexact_sample_desc": "%d sequences (every length bucket, the 8 longest) vs the reference's bench2/zu.c "
                                "(tests/golden/zuker_batch4096_sample.json), MFE and full trace" % n_exact,
        "graph_replay_equals_eager": graph_same,
        "roofline": roofline, "cpu_baseline": cpu, "served": served,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
