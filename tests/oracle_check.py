"""The timed CPU baseline of bench.py: the oracle on the batch's sequences.
Test infrastructure (uses oracle/build/liboracle.so)."""
import os
import time

import oracle_lib as o

def cpu_baseline(seqs, budget_s, threads=None):
    """The oracle (the reference loop order, bench2/zu.c:39-47) on the batch's
    sequences, one instance per host thread (ctypes releases the GIL during
    each call), taking sequences in batch order until budget_s of wall time.
    threads defaults to the host's CPU count, at most 16."""
    import threading
    nthr = threads or min(16, os.cpu_count() or 1)
    lock = threading.Lock()
    state = {"next": 0, "cells": 0, "done": 0, "busy": 0.0}
    t_end = time.perf_counter() + budget_s

    def worker():
        while True:
            with lock:
                q = state["next"]
                if q >= len(seqs) or time.perf_counter() >= t_end:
                    return
                state["next"] = q + 1
            c0 = time.perf_counter()
            o.zuker(seqs[q])
            dt = time.perf_counter() - c0
            with lock:
                state["cells"] += (len(seqs[q]) + 1) * (len(seqs[q]) + 2) // 2
                state["done"] += 1
                state["busy"] += dt

    t0 = time.perf_counter()
    ths = [threading.Thread(target=worker) for _ in range(nthr)]
    for t in ths:
        t.start()
    for t in ths:
        t.join()
    wall = time.perf_counter() - t0
    # aggregate rate of the nthr concurrent threads: cells per second of one
    # thread's calls (timed under the concurrent load) times nthr, so the
    # uneven tail of the last sequences does not count as idle time
    rate = nthr * state["cells"] / state["busy"]
    return {"value": round(rate / 1e9, 6), "unit": "GCUPS", "cores": nthr, "kind": "port",
            "sample": "first %d of the %d batch sequences in batch order (%d cells, MFE + backtrace), "
                      "%d threads running one sequence each concurrently, %.1f s of calls, %.1f s wall" % (
                          state["done"], len(seqs), state["cells"], nthr, state["busy"], wall),
            "host_threads_available": os.cpu_count()}
