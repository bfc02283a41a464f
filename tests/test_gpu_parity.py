"""GPU parity: the HIP path through the C ABI against the reference's golden
vectors and the oracle, bit-exact (integer scores, full backtraces)."""
import json
import os
import random

import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ZU_TL = [2] * 6 + [1] * 9


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def _zn(tr):
    return [[a, b, r, p0, (p1 if 0 <= r < 15 and ZU_TL[r] > 1 else 0)] for a, b, r, p0, p1 in tr]


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


ZU = _load("zuker_ref.json")
NU = _load("nussinov_ref.json")


@pytest.mark.parametrize("fx", NU, ids=lambda f: "n%d" % len(f["seq"]))
def test_nussinov_reference(ctx, fx):
    print("nussinov n=%d" % len(fx["seq"]), flush=True)
    score, tr = ctx.nussinov(fx["seq"])
    assert score == fx["score"]
    assert [[a, b, r, p0, 0] for a, b, r, p0, _ in tr] == fx["trace"]


@pytest.mark.parametrize("fx", ZU, ids=lambda f: "n%d" % len(f["seq"]))
def test_zuker_reference(ctx, fx):
    print("zuker n=%d" % len(fx["seq"]), flush=True)
    score, tr = ctx.zuker(fx["seq"])
    assert score == fx["score"]
    assert _zn(tr) == fx["trace"]


def test_zuker_batch_equals_alone(ctx):
    """Every result of a mixed batch equals the reference, i.e. the same as alone."""
    seqs = [f["seq"] for f in ZU]
    res, trs = ctx.zuker_batch(seqs)
    for f, s, t in zip(ZU, res, trs):
        assert s == f["score"] and _zn(t) == f["trace"], len(f["seq"])


def test_zuker_batch_random_vs_oracle(ctx):
    from dpx import gen
    seqs = gen.rna_batch(40, 5, 400, 77)
    res, trs = ctx.zuker_batch(seqs)
    for s, r, t in zip(seqs, res, trs):
        osc, otr = o.zuker(s)
        assert r == osc and _zn(t) == _zn(otr), len(s)
    # parse only (no trace buffer) gives the same energies
    res2, _ = ctx.zuker_batch(seqs, trace=False)
    assert res2 == res


@pytest.mark.parametrize("m", [0, 1, 3])
def test_nussinov_min_loop_vs_oracle(ctx, m):
    from dpx import gen
    seqs = [gen.rna_uniform(n, 900 + n + m) for n in [0, 1, 2, 5, 31, 32, 33, 64, 65, 200, 511]]
    res, trs = ctx.nussinov_batch(seqs, min_loop=m)
    for s, r, t in zip(seqs, res, trs):
        osc, otr = o.nussinov(s, min_loop=m)
        assert r == osc and [x[:4] for x in t] == [x[:4] for x in otr], (len(s), m)


def test_matchain_reference_values(ctx):
    # Tests.scala:16-25
    assert ctx.matchain([10, 100, 5], [100, 5, 50])[0] == 7500
    assert ctx.matchain([1, 2, 20, 2, 4, 2, 1, 7], [2, 20, 2, 4, 2, 1, 7, 3])[0] == 122


@pytest.mark.parametrize("n", [1, 2, 3, 31, 32, 33, 64, 100, 257, 512])
def test_matchain_vs_oracle(ctx, n):
    from dpx import gen
    rows, cols = gen.matrix_chain(n, 40 + n)
    assert ctx.matchain(rows, cols) == o.matchain(rows, cols)


def test_matchain_wraps_like_jvm(ctx):
    rows, cols = [2000, 60000, 30000], [60000, 30000, 50000]
    assert ctx.matchain(rows, cols) == o.matchain(rows, cols)


SW_SIZES = [(0, 0), (0, 7), (7, 0), (1, 1), (5, 9), (63, 64), (64, 63), (65, 129), (200, 77), (333, 1000)]


@pytest.mark.parametrize("n1,n2", SW_SIZES)
@pytest.mark.parametrize("sc", [(10, -3, -3, -1), (2, -3, -5, -2)])
def test_swat_vs_oracle(ctx, n1, n2, sc):
    from dpx import gen
    a, b = gen.dna_pair(n1, n2, n1 * 1000 + n2)
    for tf in (0, 1):
        s, best, tr = ctx.swat(a, b, *sc, trace_from=tf)
        os_, obest, otr = o.swat(a, b, *sc, from_best=bool(tf))
        assert (s, best) == (os_, obest)
        assert [x[:3] for x in tr] == [x[:3] for x in otr]


def test_swat_related_sequences(ctx):
    """A pair with a long shared segment: large local score, long trace."""
    from dpx import gen
    base = gen.dna_uniform(3000, 5)
    a = base[:2000]
    b = gen.dna_uniform(500, 6) + base[700:1900] + gen.dna_uniform(300, 7)
    for sc in [(10, -3, -3, -1), (2, -3, -5, -2)]:
        for tf in (0, 1):
            s, best, tr = ctx.swat(a, b, *sc, trace_from=tf)
            os_, obest, otr = o.swat(a, b, *sc, from_best=bool(tf))
            assert (s, best) == (os_, obest) and [x[:3] for x in tr] == [x[:3] for x in otr]


def test_swat_full_size_vs_oracle(ctx):
    """The 16384 x 16384 alignment workload (seed 1), against the oracle."""
    from dpx import gen
    a, b = gen.dna_pair(16384, 16384, 1)
    s, best, tr = ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
    os_, obest, otr = o.swat(a, b, 2, -3, -5, -2, from_best=True)
    assert (s, best) == (os_, obest)
    assert [x[:3] for x in tr] == [x[:3] for x in otr]


def test_zuker_longest_sequence(ctx):
    """The longest sequence dpx_zuker accepts (14000 nt: the finalize keeps
    the st column, the sequence and the energy tables' head in LDS) alone,
    in a batch beside a short one and through RNAfold; one more base is
    refused."""
    import dpx
    from dpx import gen
    s = gen.rna_gc(14000, 11)
    print("zuker n=14000", flush=True)
    score, tr = ctx.zuker(s)
    assert tr and all(0 <= t[0] <= t[1] <= len(s) for t in tr)
    res, trs = ctx.zuker_batch([s, s[:500]])
    assert res[0] == score and trs[0] == tr
    alone = ctx.zuker(s[:500])
    assert res[1] == alone[0] and trs[1] == alone[1]
    print("rnafold n=14000", flush=True)
    rs, rtr = ctx.rnafold(s)
    assert ctx.rnafold_batch([s])[0][0] == rs
    with pytest.raises(dpx.DpxError) as e:
        ctx.zuker(s + "A")
    assert e.value.code == dpx.EINVAL


def test_errors(ctx):
    import dpx
    with pytest.raises(dpx.DpxError) as e:
        ctx.zuker("ACGUN")
    assert e.value.code == dpx.EINVAL
    with pytest.raises(dpx.DpxError) as e:
        ctx.nussinov("ACGU", min_loop=-1)
    assert e.value.code == dpx.EINVAL
    bare = dpx.Context(rna_params=None)
    with pytest.raises(dpx.DpxError) as e:
        bare.zuker("GGGGAAAACCCC")
    assert e.value.code == dpx.ENOPARAMS
    with pytest.raises(dpx.DpxError) as e:
        bare.set_rna_params("/nonexistent.par")
    assert e.value.code == dpx.EPARAMS
    bare.close()
    # a context still works after errors
    assert ctx.nussinov("GGGAAACCC")[0] == o.nussinov("GGGAAACCC")[0]


def test_trace_buffer_too_small(ctx):
    import ctypes
    import dpx
    seq = ZU[3]["seq"]
    tr = (dpx.Trace * 2)()
    score, tl = ctypes.c_int32(), ctypes.c_int()
    rc = ctx._L.dpx_zuker(ctx.h, seq.encode(), len(seq), ctypes.byref(score), tr, 2, ctypes.byref(tl))
    assert rc == dpx.ETRACE and tl.value == len(ZU[3]["trace"])
    assert score.value == ZU[3]["score"]


def test_split_middle_product_equals_unsplit(ctx):
    """A launch of at most 8 instances splits the middle product of each tile
    at least 9 blocks off the diagonal into part tasks whose partial results
    the last one to finish reduces; a launch of 9 or more does not.  The same
    sequences give the same results either way, and the oracle's."""
    from dpx import gen
    longs = [gen.rna_gc(n, 700 + n) for n in (290, 700, 1500)]
    fill = [gen.rna_gc(40 + k, 800 + k) for k in range(6)]
    for name, fn in (("zuker", ctx.zuker_batch), ("rnafold", ctx.rnafold_batch),
                     ("nussinov", lambda s: ctx.nussinov_batch(s, min_loop=3))):
        print("split vs unsplit: %s" % name, flush=True)
        split = fn(longs)
        whole = fn(longs + fill)
        assert split[0] == whole[0][:3] and split[1] == whole[1][:3], name
        if name == "zuker":
            osc, otr = o.zuker(longs[0])
            assert split[0][0] == osc and _zn(split[1][0]) == _zn(otr)
        if name == "nussinov":
            osc, otr = o.nussinov(longs[1], min_loop=3)
            assert split[0][1] == osc and [t[:4] for t in split[1][1]] == [t[:4] for t in otr]
    rows, cols = gen.matrix_chain(600, 5)
    print("split: matrix chain n=600 vs oracle", flush=True)
    assert ctx.matchain(rows, cols) == o.matchain(rows, cols)


def test_plan_graph_replays(ctx):
    """A batch captured as a HIP graph: every replay, with eager work of the
    same context in between, gives the eager batch's results."""
    import dpx
    from dpx import gen
    seqs = gen.rna_batch(60, 5, 600, 91)
    want = ctx.zuker_batch(seqs)
    with ctx.plan(seqs) as p:
        for rep in range(4):
            ms = p.run()
            assert ms > 0
            f, z = p.timing()  # events between the fill and finalize graphs
            assert f > 0 and z > 0 and f + z <= ms * 1.05 + 0.05, (f, z, ms)
            assert p.results() == want, rep
            ctx.nussinov("GGGGAAAACCCC")  # eager work between replays
    nus = gen.rna_batch(30, 1, 300, 92)
    want_n = ctx.nussinov_batch(nus, min_loop=3)
    with ctx.plan(nus, kind="nussinov", min_loop=3) as p:
        for rep in range(3):
            p.run()
            assert p.results() == want_n, rep
    # energy parameters are pinned while a plan exists
    p = ctx.plan(seqs[:3], trace=False)
    with pytest.raises(dpx.DpxError) as e:
        ctx.set_rna_params(dpx.PAR_TURNER2004)
    assert e.value.code == dpx.EBUSY
    p.run()
    assert p.results()[0] == want[0][:3]
    p.close()
