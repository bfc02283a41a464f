"""Writes a launch-shape table with dpx_autotune on the current device:

    python -m dpx.tune [--out PATH] [--kinds zuker,swat,...] [--reps N]

Every kind is tuned over all length buckets (dpx_autotune checks that every
candidate shape returns identical results); kinds not listed keep the rows
of the existing table at --out.  The table is what dpx.Context loads by
default (tuning/mi355x.json)."""
import argparse
import json
import os
import sys
import time

if __package__ in (None, ""):
    sys.path.insert(0, os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
import dpx  # noqa: E402


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=dpx.TUNING_MI355X)
    ap.add_argument("--kinds", default="zuker,rnafold,nussinov,matchain,swat")
    ap.add_argument("--reps", type=int, default=3)
    args = ap.parse_args()
    old = {}
    if os.path.exists(args.out):
        with open(args.out) as f:
            old = json.load(f)
    ctx = dpx.Context(tuning=args.out if os.path.exists(args.out) else None)
    report = dict(old.get("report", {}))
    for kind in args.kinds.split(","):
        t0 = time.perf_counter()
        report[kind] = ctx.autotune(kind, reps=args.reps)
        chosen = {r["bucket"]: (r["wg_per_cu"], r["chunk"]) for r in report[kind] if r["chosen"]}
        print("%s: %.1f s, chosen (wg_per_cu, chunk) by bucket %s" % (kind, time.perf_counter() - t0, chosen), flush=True)
    out = {"device": "MI355X (gfx950)", "reps": args.reps, "table": ctx.tuning_table(), "report": report,
           "note": "launch shapes chosen by dpx_autotune on one MI355X (best of %d device timings per candidate on "
                   "seeded synthetic launches of each length bucket); loaded by dpx.Context by default" % args.reps}
    with open(args.out, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", args.out)


if __name__ == "__main__":
    main()
