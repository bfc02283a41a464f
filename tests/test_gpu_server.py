"""Serving loop (continuous batching) through the C ABI: every result of a
mixed stream equals the same request solved alone, bit-exact, including
backtraces; admission against the HBM budget."""
import random
import time

import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def _mixed_requests(seed):
    from dpx import gen
    rnd = random.Random(seed)
    reqs = []
    rna = gen.rna_batch(48, 20, 700, seed, gc=0.5)
    for q, s in enumerate(rna):
        reqs.append(("zuker", s))
        if q % 6 == 0:
            reqs.append(("nussinov", s[: len(s) // 2 + 1]))
    for q in range(8):
        n1, n2 = rnd.randint(0, 900), rnd.randint(0, 900)
        a, b = gen.dna_pair(n1, n2, seed * 100 + q)
        reqs.append(("swat", a, b, (2, -3, -5, -2) if q % 2 else (10, -3, -3, -1), q % 2))
    for q in range(6):
        rows, cols = gen.matrix_chain(rnd.randint(1, 300), seed * 10 + q)
        reqs.append(("matchain", rows, cols))
    rnd.shuffle(reqs)
    return reqs


def _alone(ctx, r):
    if r[0] == "zuker":
        return ctx.zuker(r[1])
    if r[0] == "nussinov":
        return ctx.nussinov(r[1])
    if r[0] == "matchain":
        return ctx.matchain(r[1], r[2])
    s, best, tr = ctx.swat(r[1], r[2], *r[3], trace_from=r[4])
    return (s, best), tr


def _submit(srv, r):
    if r[0] == "zuker":
        return srv.submit_zuker(r[1])
    if r[0] == "nussinov":
        return srv.submit_nussinov(r[1])
    if r[0] == "matchain":
        return srv.submit_matchain(r[1], r[2])
    return srv.submit_swat(r[1], r[2], *r[3], trace_from=r[4])


def _served(res, r):
    if r[0] == "swat":
        return (res.score, (res.best, res.best_i, res.best_j))
    return res.score


@pytest.mark.parametrize("slots,budget", [(3, 4 << 30), (1, 4 << 30), (2, 96 << 20)])
def test_mixed_stream_same_as_alone(ctx, slots, budget):
    import dpx
    print("server slots=%d budget=%d" % (slots, budget), flush=True)
    reqs = _mixed_requests(11 + slots)
    alone = [_alone(ctx, r) for r in reqs]
    with dpx.Server(ctx, hbm_budget=budget, slots=slots, max_batch=64) as srv:
        ids = []
        for q, r in enumerate(reqs):
            ids.append(_submit(srv, r))
            if q % 10 == 9:
                time.sleep(0.002)  # arrivals while earlier launches run
        got = srv.drain(ids)
        st = srv.stats()
    for r, rid, want in zip(reqs, ids, alone):
        res, tr = got[rid]
        assert res.status == 0, (r[0], res.status)
        assert _served(res, r) == want[0], r[0]
        assert tr == want[1], r[0]
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["peak_bytes_in_flight"] <= budget
    assert st["bytes_in_flight"] == 0
    assert st["max_batch_seen"] > 1  # requests were batched


def test_served_zuker_vs_oracle(ctx):
    import dpx
    from dpx import gen
    seqs = gen.rna_batch(24, 5, 300, 4242, gc=0.5)
    with dpx.Server(ctx, slots=2, max_batch=16) as srv:
        ids = [srv.submit_zuker(s) for s in seqs]
        got = srv.drain(ids)
    for s, rid in zip(seqs, ids):
        res, tr = got[rid]
        osc, otr = o.zuker(s)
        assert res.score == osc
        assert [t[:4] for t in tr] == [t[:4] for t in otr]


def test_admission_rejects_oversize(ctx):
    import dpx
    from dpx import gen
    with dpx.Server(ctx, hbm_budget=8 << 20, slots=1, max_batch=8) as srv:
        with pytest.raises(dpx.DpxError) as e:
            srv.submit_zuker(gen.rna_gc(2000, 1, 0.5))
        assert e.value.code == dpx.ENOMEM
        ok = srv.submit_zuker("GGGGAAAACCCC")
        got = srv.drain([ok])
        assert got[ok][0].score == ctx.zuker("GGGGAAAACCCC")[0]
        assert srv.stats()["rejected"] == 1


def test_submit_errors(ctx):
    import dpx
    with dpx.Server(ctx, slots=1) as srv:
        with pytest.raises(dpx.DpxError) as e:
            srv.submit_zuker("ACGUN")
        assert e.value.code == dpx.EINVAL
    srv2 = dpx.Server(ctx, slots=1)
    try:
        with pytest.raises(dpx.DpxError) as e:
            ctx.set_rna_params(dpx.PAR_TURNER2004)
        assert e.value.code == dpx.EBUSY
    finally:
        srv2.close()
    ctx.set_rna_params(dpx.PAR_TURNER2004)  # allowed again once no server exists


def test_alignments_batched_in_one_launch(ctx):
    """Smith-Waterman requests that wait together run as one launch over the
    bands of all of them (the alignment of each band from a table); every
    result and trace equals the same pair aligned alone, also with empty
    sequences, mixed scores and both trace starts; traced and untraced
    requests never share a launch."""
    import dpx
    from dpx import gen
    rnd = random.Random(77)
    reqs = []
    for q in range(40):
        n1, n2 = rnd.randint(200, 500), rnd.randint(200, 500)
        if q % 13 == 5:
            n1 = 0
        if q % 17 == 3:
            n2 = 0
        a, b = gen.dna_pair(n1, n2, 900 + q)
        sc = (2, -3, -5, -2) if q % 3 else (10, -3, -3, -1)
        reqs.append((a, b, sc, q % 2, q % 4 != 3))
    print("batched alignments: %d requests" % len(reqs), flush=True)
    with dpx.Server(ctx, hbm_budget=4 << 30, slots=1, max_batch=64) as srv:
        ids = [srv.submit_swat(a, b, *sc, trace_from=tf, trace=tr) for a, b, sc, tf, tr in reqs]
        got = srv.drain(ids)
        st = srv.stats()
    assert st["completed"] == len(reqs) and st["failed"] == 0
    assert st["max_batch_seen"] > 8 and st["launches"] < len(reqs) // 2
    for (a, b, sc, tf, trace), rid in zip(reqs, ids):
        res, tr = got[rid]
        assert res.status == 0
        s, best, want_tr = ctx.swat(a, b, *sc, trace_from=tf)
        assert (res.score, (res.best, res.best_i, res.best_j)) == (s, best), (len(a), len(b), sc, tf)
        assert tr == (want_tr if trace else []), (len(a), len(b), sc, tf)
    # and two of them against the oracle
    for (a, b, sc, tf, trace), rid in list(zip(reqs, ids))[:2]:
        s, best, otr = o.swat(a, b, *sc, from_best=bool(tf))
        res, tr = got[rid]
        assert (res.score, (res.best, res.best_i, res.best_j)) == (s, best)
        assert [t[:3] for t in tr] == ([t[:3] for t in otr] if trace else [])
