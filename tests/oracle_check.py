"""Oracle-side helpers for bench.py and smoke(): the checker and the timed
CPU baseline.  Test infrastructure (uses oracle/build/liboracle.so)."""
import os
import time

import oracle_lib as o

ZU_TL = [2] * 6 + [1] * 9


def _zn(tr):
    return [(a, b, r, p0, (p1 if 0 <= r < 15 and ZU_TL[r] > 1 else 0)) for a, b, r, p0, p1 in tr]


def zuker_sample(seqs, results, traces, k=8):
    """Bit-exact check of k instances of a batch (the k shortest) against the oracle."""
    idx = sorted(range(len(seqs)), key=lambda q: len(seqs[q]))[:k]
    for q in idx:
        sc, tr = o.zuker(seqs[q])
        if sc != results[q] or (traces is not None and _zn(tr) != _zn(traces[q])):
            return False
    return True


def cpu_baseline(seqs, budget_s):
    """Single-thread oracle (the reference loop order, bench2/zu.c:39-47) on
    the batch's sequences in order until budget_s of CPU time."""
    done_cells, t_total, k = 0, 0.0, 0
    for s in seqs:
        t0 = time.process_time()
        o.zuker(s)
        t_total += time.process_time() - t0
        done_cells += (len(s) + 1) * (len(s) + 2) // 2
        k += 1
        if t_total >= budget_s:
            break
    return {"value": round(done_cells / t_total / 1e9, 6), "unit": "GCUPS", "cores": 1, "kind": "port",
            "sample": "first %d of the %d batch sequences (%d cells, MFE + backtrace), %.1f s" % (k, len(seqs), done_cells, t_total),
            "host_threads_available": os.cpu_count()}
