#!/usr/bin/env python3
"""Benchmark of the MI355X DP backend.

Workload (one "step"): batched Zuker MFE + backtrace of 4096 RNA sequences,
lengths log-uniform in [50, 2000], i.i.d. 50% GC (seed 3), one persistent
tile-queue launch plus one finalize launch.  Prints one JSON line.

value = cell updates per second of the batched step, with inputs resident in
HBM: the device time of the fill and finalize kernels (HIP events on the
library's stream).  Also reported: the same sequences as plain per-instance
launches (a fixed 1-in-16 sample), the end-to-end rate including PCIe copies,
the roofline of the fill kernel and a single-thread CPU baseline (the oracle,
a bounded sample of the same sequences).
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "lamp-dp-mt_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

METRIC = "GCUPS batched/served vs plain per-instance (bit-exact); p50/p99 fold latency"
PEAK_INT32_TOPS = 256 * 64 * 2.4e9 / 1e12  # 256 CUs x 64 lanes x 2.4 GHz


def cells(n):
    return (n + 1) * (n + 2) // 2


def x_candidates(n):
    """Split points of the ml/ml1 min-plus term, k in [i+7, j-7]: sum over
    diagonals d=j-i of (n-d+1)*(d-13)."""
    m = n - 13
    return m * (m + 1) * (m + 2) // 6 if m > 0 else 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--plain-stride", type=int, default=16)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(instances shard across GPUs without communication, see DESIGN.md)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import dpx
    from dpx import gen

    seqs = gen.rna_batch(args.count, 50, 2000, 3, gc=0.5)
    total_cells = sum(cells(len(s)) for s in seqs)
    total_x = sum(x_candidates(len(s)) for s in seqs)
    print("bench: %d sequences, %.3e cells" % (len(seqs), total_cells), file=sys.stderr, flush=True)

    ctx = dpx.Context()
    for _ in range(args.warmup):
        ctx.zuker_batch(seqs)
    fills, fins, walls = [], [], []
    results = traces = None
    for _ in range(args.steps):
        t0 = time.perf_counter()
        results, traces = ctx.zuker_batch(seqs)
        walls.append(time.perf_counter() - t0)
        f, z = ctx.last_timing()
        fills.append(f)
        fins.append(z)
        print("bench: step fill=%.2fms finalize=%.2fms wall=%.2fms" % (f, z, walls[-1] * 1e3), file=sys.stderr, flush=True)
    dev_ms = (sum(fills) + sum(fins)) / args.steps
    value = total_cells / (dev_ms * 1e-3) / 1e9
    wall_ms = sum(walls) / args.steps * 1e3

    # plain per-instance launches over a fixed sample of the same sequences
    sample = seqs[::args.plain_stride]
    p_cells, p_ms = 0, 0.0
    for s in sample:
        r, _ = ctx.zuker(s)
        f, z = ctx.last_timing()
        p_cells += cells(len(s))
        p_ms += f + z
    plain = p_cells / (p_ms * 1e-3) / 1e9
    print("bench: plain per-instance %.3f GCUPS over %d sequences" % (plain, len(sample)), file=sys.stderr, flush=True)

    # bit-exact spot check against the oracle on the shortest sequences of the batch
    import oracle_check
    exact = oracle_check.zuker_sample(seqs, results, traces, k=8)

    fill_ms = sum(fills) / args.steps
    achieved = 2.0 * total_x / (fill_ms * 1e-3) / 1e12
    roofline = {"bound": "valu", "achieved": round(achieved, 4), "peak": round(PEAK_INT32_TOPS, 2), "unit": "TOP/s",
                "frac": round(achieved / PEAK_INT32_TOPS, 5), "traffic": None,
                "kernel": "tri_fill<GZu>", "ops": "2 int32 ops (add, min) per ml/ml1 split candidate"}

    cpu = None
    if not args.no_cpu:
        cpu = oracle_check.cpu_baseline(seqs, args.cpu_seconds)

    line = {
        "metric": METRIC, "value": round(value, 4), "unit": "GCUPS", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(dev_ms, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "int32", "data": "synthetic",
        "config": {"workload": "batched Zuker MFE + backtrace: %d RNA sequences, lengths log-uniform [50,2000], "
                               "50%% GC, seed 3, Turner 2004 parameters" % len(seqs),
                   "instances": len(seqs), "cells": total_cells, "parallelism": "single GPU, persistent tile queue"},
        "plain_per_instance_gcups": round(plain, 4), "speedup_vs_plain": round(value / plain, 2),
        "wall_ms_per_step_pcie": round(wall_ms, 3), "fill_ms": round(fill_ms, 3),
        "finalize_ms": round(sum(fins) / args.steps, 3), "bitexact_sample": exact,
        "roofline": roofline, "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
