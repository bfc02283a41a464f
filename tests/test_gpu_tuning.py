"""Launch-shape tuning: every shape the tuning table can select gives the
same results (bit-exact against the oracle), and the autotuner times the
candidates of each length bucket, keeps the fastest and reports it."""
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def test_bucket_edges():
    import dpx
    L = dpx.lib()
    assert [L.dpx_tune_bucket(n) for n in (0, 1, 64, 65, 128, 129, 4096, 4097, 10 ** 6)] == [0, 0, 0, 1, 1, 2, 6, 7, 7]
    assert L.dpx_tune_bucket(-1) == -1


def test_set_get_and_errors(ctx):
    import dpx
    ctx.set_tuning("zuker", 3, wg_per_cu=2)
    assert ctx.get_tuning("zuker", 3) == {"wg_per_cu": 2, "chunk": 0}
    ctx.set_tuning("zuker", 3)
    for bad in [dict(kind="zuker", bucket=8), dict(kind="zuker", bucket=-1), dict(kind=9, bucket=0),
                dict(kind="swat", bucket=0, chunk=24), dict(kind="swat", bucket=0, wg_per_cu=-1)]:
        with pytest.raises(dpx.DpxError) as e:
            ctx.set_tuning(bad["kind"], bad["bucket"], bad.get("wg_per_cu", 0), bad.get("chunk", 0))
        assert e.value.code == dpx.EINVAL
    table = ctx.tuning_table()
    assert set(table) == {"nussinov", "zuker", "swat", "matchain", "rnafold"}
    ctx.load_tuning_table(table)


@pytest.mark.parametrize("w", [1, 2, 3])
def test_zuker_batch_every_grid(ctx, w):
    from dpx import gen
    seqs = gen.rna_batch(16, 20, 500, 5)
    for b in range(8):
        ctx.set_tuning("zuker", b, wg_per_cu=w)
    try:
        res, trs = ctx.zuker_batch(seqs)
    finally:
        for b in range(8):
            ctx.set_tuning("zuker", b)
    for s, r, t in zip(seqs, res, trs):
        osc, otr = o.zuker(s)
        assert r == osc and [x[:4] for x in t] == [x[:4] for x in otr], len(s)


@pytest.mark.parametrize("chunk,wg", [(16, 0), (64, 0), (16, 1), (64, 1)])
def test_swat_every_chunk(ctx, chunk, wg):
    """Chunk sizes, and one band per CU (more bands than workgroups: the
    persistent workgroups take further bands in order)."""
    from dpx import gen
    for b in range(8):
        ctx.set_tuning("swat", b, wg_per_cu=wg, chunk=chunk)
    try:
        for n1, n2, seed in [(1, 1, 1), (63, 65, 2), (130, 77, 3), (1000, 999, 4), (17000, 300, 5)]:
            a, bb = gen.dna_pair(n1, n2, seed)
            print("swat chunk=%d wg=%d %dx%d" % (chunk, wg, n1, n2), flush=True)
            got = ctx.swat(a, bb, 2, -3, -5, -2, trace_from=1)
            want = o.swat(a, bb, 2, -3, -5, -2, from_best=True)
            assert got[0] == want[0] and got[1] == want[1]
            assert [t[:3] for t in got[2]] == [t[:3] for t in want[2]]
    finally:
        for b in range(8):
            ctx.set_tuning("swat", b)


@pytest.mark.parametrize("kind", ["nussinov", "zuker", "rnafold", "matchain", "swat"])
def test_autotune_reports_and_keeps_fastest(ctx, kind):
    print("autotune %s" % kind, flush=True)
    rep = ctx.autotune(kind, buckets=range(0, 3), reps=1)
    assert rep
    for b in range(0, 3):
        rows = [r for r in rep if r["bucket"] == b]
        assert len(rows) >= 1 and sum(r["chosen"] for r in rows) == 1
        best = min(rows, key=lambda r: r["ms"])
        chosen = [r for r in rows if r["chosen"]][0]
        assert chosen["ms"] == best["ms"]
        assert ctx.get_tuning(kind, b) == {"wg_per_cu": chosen["wg_per_cu"], "chunk": chosen["chunk"]}
        assert all(r["count"] >= 1 and r["n_lo"] <= r["n_hi"] for r in rows)
    for b in range(0, 3):
        ctx.set_tuning(kind, b)


def test_zuker_launch_shapes_agree_repeatedly():
    """Regression: the Zuker finalize once read st(n) before thread 0 had
    written it (an LDS race, visible as launches of the same batch disagreeing
    now and then).  Every autotune compares all launch shapes bit for bit."""
    import dpx
    for it in range(12):
        c = dpx.Context()
        try:
            c.autotune("zuker", buckets=range(0, 2), reps=1)
        finally:
            c.close()
