"""GPU parity of the RNAfold grammar (v4/examples/RNAFold.scala:20-64):
minimum free energy and the full backtrace of the HIP path, through the C
ABI, bit-exact against the oracle restatement (itself pinned by brute force
over all structures in test_oracle_rnafold.py)."""
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def _check(ctx, s):
    score, tr = ctx.rnafold(s)
    osc, otr = o.rnafold(s)
    assert score == osc, (len(s), score, osc)
    assert tr == otr, len(s)


@pytest.mark.parametrize("n", list(range(0, 12)) + [30, 31, 32, 33, 34, 63, 64, 65, 66, 97])
def test_rnafold_small_sizes(ctx, n):
    """Tile edges (32) and the smallest foldable lengths."""
    from dpx import gen
    print("rnafold n=%d" % n, flush=True)
    for seed in range(3):
        _check(ctx, gen.rna_gc(n, 1000 + 17 * n + seed))


@pytest.mark.parametrize("n,seed", [(150, 1), (257, 2), (400, 3), (700, 4)])
def test_rnafold_vs_oracle(ctx, n, seed):
    from dpx import gen
    print("rnafold n=%d seed=%d" % (n, seed), flush=True)
    _check(ctx, gen.rna_gc(n, seed))
    _check(ctx, gen.rna_uniform(n, seed + 50))


def test_rnafold_long_loops(ctx):
    """Interior loops at the 30-base window edge and hairpins longer than 30
    (Jacobson-Stockmayer extrapolation): G/C stems around long A runs."""
    for inner, left, right in [(40, 28, 1), (35, 29, 29), (50, 1, 29), (60, 30, 30), (8, 29, 28)]:
        s = "GGGGC" + "A" * left + "GGGCC" + "A" * inner + "GGCCC" + "A" * right + "GCCCC"
        _check(ctx, s)
    _check(ctx, "GGGGG" + "A" * 200 + "CCCCC")


def test_rnafold_empty_answer(ctx):
    # no canonical pair at all: the grammar admits no structure
    assert ctx.rnafold("AAAAAAAAAA") == o.rnafold("AAAAAAAAAA") == (999999, [])
    assert ctx.rnafold("")[0] == 999999


def test_rnafold_batch_vs_oracle(ctx):
    from dpx import gen
    seqs = gen.rna_batch(48, 1, 500, 123) + ["", "A", "GC", "GAC", "AAAAAA"]
    res, trs = ctx.rnafold_batch(seqs)
    for s, r, t in zip(seqs, res, trs):
        osc, otr = o.rnafold(s)
        assert r == osc and t == otr, len(s)
    res2, _ = ctx.rnafold_batch(seqs, trace=False)
    assert res2 == res


def test_rnafold_plan_replays(ctx):
    from dpx import gen
    seqs = gen.rna_batch(24, 5, 300, 321)
    want = ctx.rnafold_batch(seqs)
    with ctx.plan(seqs, kind="rnafold") as p:
        for rep in range(3):
            p.run()
            assert p.results() == want, rep
            ctx.zuker("GGGGAAAACCCC")


def test_rnafold_served_equals_alone(ctx):
    import dpx
    from dpx import gen
    seqs = gen.rna_batch(20, 5, 300, 555)
    with dpx.Server(ctx, hbm_budget=1 << 30, slots=2, max_batch=8) as srv:
        ids = [srv.submit_rnafold(s) for s in seqs]
        ids.append(srv.submit_zuker(seqs[0]))
        got = srv.drain(ids, timeout_s=60)
    for s, rid in zip(seqs, ids):
        r, tr = got[rid]
        osc, otr = o.rnafold(s)
        assert r.status == dpx.OK and r.score == osc and tr == otr, len(s)
    r, _ = got[ids[-1]]
    assert r.score == o.zuker(seqs[0])[0]


def _large():
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rnafold_large.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("fx", _large(), ids=lambda f: "%s-n%d" % (f["gen"], f["n"]))
def test_rnafold_full_size_fixture(ctx, fx):
    """n = 2048 (the Zuker configuration's input) and n = 1500 against the
    oracle's answers committed by tests/golden/make_rnafold_large.py."""
    from dpx import gen
    seq = getattr(gen, fx["gen"])(fx["n"], fx["seed"])
    print("rnafold fixture n=%d" % len(seq), flush=True)
    score, tr = ctx.rnafold(seq)
    assert score == fx["score"]
    assert [list(t) for t in tr] == fx["trace"]
