"""The benchmark configurations at their own sizes, through the C ABI.

* Nussinov n=1024 (uniform ACGU, seed 0) with a minimum hairpin loop of 3,
  against the oracle.
* Batched Zuker: the 4096-sequence set (lengths log-uniform [50,2000], 50% GC,
  seed 3) captured as a plan (HIP graph) exactly as the benchmark runs it;
  every result equals the eager batch, and a stratified sample (every length
  bucket, the 8 longest sequences) equals the reference's own generated CPU
  code (bench2/zu.c, committed fixtures), score and full trace.
* The served mix: the first 2 s of the Poisson stream at 300 requests/s
  (Zuker 30-3000 nt, SW pairs 500-20000 bp, matrix chains of 64-512) through
  the continuous-batching server; every result equals the same request solved
  alone, the sampled Zuker requests (up to n=2960) equal the reference's
  fixtures, and SW and matrix-chain requests equal the oracle.
"""
import json
import os
from concurrent.futures import ThreadPoolExecutor

import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ZU_TL = [2] * 6 + [1] * 9
SW = (2, -3, -5, -2)


def _zn(tr):
    return [[a, b, r, p0, (p1 if 0 <= r < 15 and ZU_TL[r] > 1 else 0)] for a, b, r, p0, p1 in tr]


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def test_nussinov_1024_min_loop3_vs_oracle(ctx):
    from dpx import gen
    s = gen.rna_uniform(1024, 0)
    print("nussinov n=1024 seed 0 min_loop 3", flush=True)
    score, tr = ctx.nussinov(s, min_loop=3)
    osc, otr = o.nussinov(s, min_loop=3)
    assert score == osc
    assert [t[:4] for t in tr] == [t[:4] for t in otr]
    # the same sequence inside a batch, as a plan
    with ctx.plan([s, s[:500], s[300:]], kind="nussinov", min_loop=3) as p:
        p.run()
        res, trs = p.results()
    assert res[0] == osc and [t[:4] for t in trs[0]] == [t[:4] for t in otr]


def test_batched_zuker_4096_plan(ctx):
    from dpx import gen
    seqs = gen.rna_batch(4096, 50, 2000, 3, gc=0.5)
    fx = _load("zuker_batch4096_sample.json")
    print("batched zuker 4096 seqs seed 3: plan vs eager vs reference sample (%d)" % len(fx), flush=True)
    want = ctx.zuker_batch(seqs)
    with ctx.plan(seqs) as p:
        for rep in range(2):
            p.run()
            got = p.results()
            assert got[0] == want[0], rep
            assert got[1] == want[1], rep
    res, trs = want
    longest = sorted(range(len(seqs)), key=lambda q: -len(seqs[q]))[:8]
    sampled = {f["index"] for f in fx}
    assert set(longest) <= sampled
    for f in fx:
        q = f["index"]
        assert seqs[q] == f["seq"], q
        assert res[q] == f["score"], (q, len(seqs[q]))
        assert _zn(trs[q]) == f["trace"], (q, len(seqs[q]))
    # and a few more of every size against the live oracle (threads: ctypes
    # releases the GIL), shortest ones of each bucket
    extra = [q for q in range(0, 4096, 257) if len(seqs[q]) <= 900]
    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(lambda q: o.zuker(seqs[q]), extra))
    for q, (osc, otr) in zip(extra, outs):
        assert res[q] == osc and _zn(trs[q]) == _zn(otr), (q, len(seqs[q]))


def _alone(ctx, kind, x):
    if kind == "zuker":
        return ctx.zuker(x)
    if kind == "matchain":
        return ctx.matchain(x[0], x[1])
    s, best, tr = ctx.swat(x[0], x[1], *SW, trace_from=1)
    return (s, best), tr


def test_served_mix_real_sizes(ctx):
    import dpx
    from dpx import gen
    reqs = gen.served_stream(300.0, 2.0, 5)
    fx = {f["index"]: f for f in _load("zuker_served_sample.json")}
    print("served stream 2 s at 300/s: %d requests (%d reference-pinned Zuker)" % (len(reqs), len(fx)), flush=True)
    with dpx.Server(ctx, hbm_budget=16 << 30, slots=3, max_batch=256) as srv:
        t0 = dpx.now()
        ids = []
        for at, kind, x in reqs:
            d = t0 + at - dpx.now()
            if d > 0:
                import time
                time.sleep(d)
            if kind == "zuker":
                ids.append(srv.submit_zuker(x))
            elif kind == "swat":
                ids.append(srv.submit_swat(x[0], x[1], *SW, trace_from=1))
            else:
                ids.append(srv.submit_matchain(x[0], x[1]))
        got = srv.drain(ids, timeout_s=100)
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0 and st["rejected"] == 0
    print("served: %d launches, peak %d bytes; checking each request alone" % (st["launches"], st["peak_bytes_in_flight"]),
          flush=True)
    for q, ((at, kind, x), rid) in enumerate(zip(reqs, ids)):
        res, tr = got[rid]
        assert res.status == 0, (q, kind)
        want = _alone(ctx, kind, x)
        served = (res.score, (res.best, res.best_i, res.best_j)) if kind == "swat" else res.score
        assert served == want[0], (q, kind)
        assert tr == want[1], (q, kind)
        if q in fx:
            assert kind == "zuker" and x == fx[q]["seq"], q
            assert res.score == fx[q]["score"] and _zn(tr) == fx[q]["trace"], (q, len(x))
    # SW (shortest, median and longest by cells) and every matrix chain against the oracle
    swq = sorted([q for q, r in enumerate(reqs) if r[1] == "swat"], key=lambda q: len(reqs[q][2][0]) * len(reqs[q][2][1]))
    pick = sorted({swq[0], swq[len(swq) // 2], swq[-1]})
    chq = [q for q, r in enumerate(reqs) if r[1] == "matchain"]

    def oracle(q):
        kind, x = reqs[q][1], reqs[q][2]
        if kind == "swat":
            s, best, tr = o.swat(x[0], x[1], *SW, from_best=True)
            return (s, best), [t[:3] for t in tr]
        return o.matchain(x[0], x[1])

    with ThreadPoolExecutor(min(16, os.cpu_count() or 1)) as ex:
        outs = list(ex.map(oracle, pick + chq))
    for q, want in zip(pick + chq, outs):
        res, tr = got[ids[q]]
        if reqs[q][1] == "swat":
            assert ((res.score, (res.best, res.best_i, res.best_j)), [t[:3] for t in tr]) == want, q
        else:
            assert (res.score, tr) == want, q
