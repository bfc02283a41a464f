"""Generic ADP grammars on the GPU (dpx_grammar_*): every example grammar of
the reference compiled to HIP at run time and checked against the oracle
(score, found flag and full trace), and against the built-in kernels where
the grammar is theirs: Nussinov.scala vs dpx_nussinov (and bench2/nu.c's
golden vectors), SWatAffine.scala vs dpx_swat, MatrixMult2.scala vs
dpx_matchain."""
import json
import random

import pytest

import adp_grammars as G
import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def _vs_oracle(cg, g, in1, in2=None):
    ans, found, tr = cg.parse(in1, in2)
    _, (oans, ofound, otr) = G.run_oracle(g, in1, in2)
    assert found == bool(ofound)
    if found:
        assert ans == oans
        assert [tuple(t) for t in tr] == [tuple(t) for t in otr]
    return ans, found, tr


def test_nussinov_grammar_vs_builtin_and_golden(ctx):
    from dpx import gen
    g = G.nussinov()
    cg = ctx.grammar(g)
    assert cg.info.n_rules == 5 and cg.info.max_pos == 1 and cg.info.loops == 1
    for case in json.load(open(o.ROOT + "/tests/golden/nussinov_ref.json")):
        print("nussinov golden n=%d" % len(case["seq"]), flush=True)
        ans, found, tr = cg.parse(case["seq"])
        assert found and ans == case["score"]
        assert [list(t) for t in tr] == case["trace"]
    for n in [0, 1, 2, 31, 64, 65, 200, 700]:
        s = gen.rna_uniform(n, n + 11)
        ans, found, tr = cg.parse(s)
        score, btr = ctx.nussinov(s, min_loop=0)
        assert ans == score and tr == btr, n
    for n in [0, 3, 17, 40]:
        _vs_oracle(cg, g, gen.rna_uniform(n, 7 * n + 1))


@pytest.mark.parametrize("sc", [(2, -3, -5, -2), (10, -3, -3, -1)])
def test_swat_affine_grammar_vs_builtin(ctx, sc):
    from dpx import gen
    g = G.swat_affine(*sc)
    cg = ctx.grammar(g)
    assert cg.info.tracks == 2 and cg.info.n_rules == 10 and cg.info.loops == 0
    for n1, n2 in [(0, 0), (0, 5), (6, 0), (1, 1), (63, 64), (200, 77), (333, 1000), (1500, 1400)]:
        print("swat grammar %dx%d" % (n1, n2), flush=True)
        a, b = gen.dna_pair(n1, n2, n1 * 1000 + n2)
        ans, found, tr = cg.parse(a, b)
        s, best, btr = ctx.swat(a, b, *sc, trace_from=0)
        assert found and ans == s
        assert [t[:3] for t in tr] == [t[:3] for t in btr]
    for n1, n2 in [(5, 9), (20, 13)]:
        a, b = gen.dna_pair(n1, n2, n1 + n2)
        _vs_oracle(cg, g, a, b)


def test_matrix_chain_grammar_vs_builtin(ctx):
    g = G.matrix_chain()
    cg = ctx.grammar(g)
    rng = random.Random(4)
    for n in [1, 2, 3, 10, 33, 64, 129, 300]:
        dims = [rng.randint(1, 100) for _ in range(n + 1)]
        mats = [(dims[q], dims[q + 1]) for q in range(n)]
        print("chain grammar n=%d" % n, flush=True)
        ans, found, tr = cg.parse(mats)
        s, btr = ctx.matchain([m[0] for m in mats], [m[1] for m in mats])
        assert found and ans == s
        assert [t[:4] for t in tr] == [t[:4] for t in btr]
    mats = [(10, 100), (100, 5), (5, 50)]
    assert cg.parse(mats)[0] == 7500


@pytest.mark.parametrize("h", ["min", "max"])
def test_elmamun_vs_oracle(ctx, h):
    g = G.elmamun(h)
    cg = ctx.grammar(g)
    rng = random.Random(1 if h == "min" else 2)
    exprs = ["1+2*3*4+5", "9", "+", "1+", ""]
    for _ in range(12):
        n = rng.randint(1, 12)
        exprs.append("".join(str(rng.randint(0, 9)) + rng.choice("+*") for _ in range(n)) + str(rng.randint(0, 9)))
    for e in exprs:
        _vs_oracle(cg, g, e)
    assert cg.parse("1+2*3*4+5")[0] == (30 if h == "min" else 81)


def test_brackets_window_vs_oracle(ctx):
    g = G.brackets(8)
    cg = ctx.grammar(g)
    assert cg.info.window == 8 and cg.info.n_tabs == 1  # the unused tabulation is stripped
    rng = random.Random(6)
    cases = ["(((3)))((2))(1)((6))((((8))))(3)((4))", "((1)(2))", "(5)", "1234", "", "((((((((9))))))))"]
    for _ in range(20):
        cases.append("".join(rng.choice("()123456789") for _ in range(rng.randint(3, 60))))
    for s in cases:
        _vs_oracle(cg, g, s)


def test_long_inputs_past_the_lds_copy(ctx):
    """One-track tiles keep the whole input in LDS up to 16 KB; longer inputs
    are read through two windows around the tile and the memory fallback:
    the windowed Brackets grammar on 17000 characters against the oracle,
    and MatrixMult2.scala at n = 2100 (8-byte elements: past the copy)
    against dpx_matchain."""
    g = G.brackets(8)
    cg = ctx.grammar(g)
    rng = random.Random(7)
    s = "".join(rng.choice("()((12345))6789") for _ in range(17000))
    print("brackets grammar n=%d" % len(s), flush=True)
    _vs_oracle(cg, g, s)
    g = G.matrix_chain()
    cg = ctx.grammar(g)
    dims = [rng.randint(1, 100) for _ in range(2101)]
    mats = [(dims[q], dims[q + 1]) for q in range(2100)]
    print("chain grammar n=2100", flush=True)
    ans, found, tr = cg.parse(mats)
    c, btr = ctx.matchain([m[0] for m in mats], [m[1] for m in mats])
    assert found and ans == c
    assert [t[:4] for t in tr] == [t[:4] for t in btr]


@pytest.mark.parametrize("alg", ["sw", "nw"])
def test_seqalign_vs_oracle(ctx, alg):
    from dpx import gen
    g = G.seqalign(alg)
    cg = ctx.grammar(g)
    for n1, n2 in [(0, 0), (0, 4), (3, 0), (1, 1), (7, 9), (25, 31), (40, 12)]:
        a, b = gen.dna_pair(n1, n2, 31 * n1 + n2)
        _vs_oracle(cg, g, a, b)


def test_trace_capacity_and_errors(ctx):
    import dpx
    g = G.nussinov()
    cg = ctx.grammar(g)
    s = "GGGAAACCCAGGGAAACCCA" * 3
    full = cg.parse(s)
    assert cg.parse(s, cap=2) == full  # the binding grows the buffer after DPX_ETRACE
    bad = G.nussinov()
    spec = bad.spec()
    with pytest.raises(dpx.DpxError):
        ctx._check(ctx._L.dpx_grammar_create(ctx.h, b"{}", ctypes_ptr()))


def ctypes_ptr():
    import ctypes
    return ctypes.byref(ctypes.c_void_p())


def test_random_grammars_vs_oracle(ctx):
    """Seeded random grammars (every construct the library accepts) on
    random inputs: answer, found flag and full trace equal the oracle's."""
    import adp_random as R
    import dpx
    done = 0
    for seed in range(48):
        g = R.single_track(seed) if seed % 3 else R.two_track(seed)
        try:
            G.oracle(g)
        except ValueError:
            with pytest.raises(dpx.DpxError):
                ctx.grammar(g)
            continue
        print("random grammar seed %d" % seed, flush=True)
        cg = ctx.grammar(g)
        rng = random.Random(1000 + seed)
        for _ in range(4):
            if g.tracks == 1:
                _vs_oracle(cg, g, R.text(rng, rng.randint(0, 24)))
            else:
                _vs_oracle(cg, g, R.text(rng, rng.randint(0, 12)), R.text(rng, rng.randint(0, 12)))
        cg.close()
        done += 1
    assert done > 30


def test_batches_equal_single_parses(ctx):
    """dpx_grammar_parse_batch: ragged batches (empty inputs included) give
    every instance's single-parse result; one track, two tracks, window."""
    from dpx import gen
    rng = random.Random(12)
    g = G.nussinov()
    cg = ctx.grammar(g)
    seqs = [gen.rna_uniform(rng.choice([0, 1, 2, 5, 31, 64, 65, 130, 257]), q) for q in range(40)]
    print("nussinov grammar batch", flush=True)
    got = cg.parse_batch(seqs)
    builtin, btr = ctx.nussinov_batch(seqs, min_loop=0)
    for q, s in enumerate(seqs):
        assert got[q] == cg.parse(s), q
        assert got[q][0] == builtin[q] and got[q][2] == btr[q], q
    sw = ctx.grammar(G.swat_affine())
    pairs = [gen.dna_pair(rng.choice([0, 1, 7, 64, 100]), rng.choice([0, 3, 65, 90]), q) for q in range(24)]
    print("swat grammar batch", flush=True)
    got = sw.parse_batch(pairs)
    for q, (a, b) in enumerate(pairs):
        assert got[q] == sw.parse(a, b), q
        s, _, btr = ctx.swat(a, b, 2, -3, -5, -2)
        assert got[q][0] == s and [t[:3] for t in got[q][2]] == [t[:3] for t in btr]
    br = ctx.grammar(G.brackets(8))
    texts = ["", "(1)", "((2))", "(((3)))((2))(1)", "12", "((((((((9))))))))"] + \
        ["".join(rng.choice("()123") for _ in range(rng.randint(0, 40))) for _ in range(10)]
    print("brackets grammar batch", flush=True)
    got = br.parse_batch(texts)
    for q, t in enumerate(texts):
        assert got[q] == br.parse(t), q
        _vs_oracle(br, G.brackets(8), t)


def test_random_grammar_batches_vs_oracle(ctx):
    import adp_random as R
    done = 0
    for seed in range(1, 40, 3):
        g = R.single_track(seed) if seed % 3 else R.two_track(seed)
        try:
            G.oracle(g)
        except ValueError:
            continue
        cg = ctx.grammar(g)
        rng = random.Random(seed)
        if g.tracks == 1:
            ins = [R.text(rng, rng.randint(0, 20)) for _ in range(7)]
            got = cg.parse_batch(ins)
            for q, x in enumerate(ins):
                _, (a, f, tr) = G.run_oracle(g, x)
                assert got[q][1] == bool(f) and (not f or (got[q][0] == a and got[q][2] == [tuple(t) for t in tr])), (seed, q)
        else:
            ins = [(R.text(rng, rng.randint(0, 10)), R.text(rng, rng.randint(0, 10))) for _ in range(7)]
            got = cg.parse_batch(ins)
            for q, (x, y) in enumerate(ins):
                _, (a, f, tr) = G.run_oracle(g, x, y)
                assert got[q][1] == bool(f) and (not f or (got[q][0] == a and got[q][2] == [tuple(t) for t in tr])), (seed, q)
        cg.close()
        done += 1
    assert done >= 8


def test_zuker_grammar_vs_reference_vectors_and_builtin(ctx):
    """Zuker.scala (ZukerGrammar + ZukerMFE, librna calls evaluated on the
    device) through the generic path: bench2/zu.c's golden vectors (the
    reference's generated code for this grammar) and dpx_zuker."""
    import dpx.examples as ex
    from dpx import gen
    g = ex.zuker()
    cg = ctx.grammar(g)
    i = cg.info
    assert (i.n_rules, i.max_pos, i.n_tabs, i.rna, i.axiom_id) == (15, 2, 4, 1, 12)  # zu.h:1-5
    fx = json.load(open(o.ROOT + "/tests/golden/zuker_ref.json"))
    for case in fx:
        print("zuker grammar golden n=%d" % len(case["seq"]), flush=True)
        ans, found, tr = cg.parse(case["seq"])
        assert found and ans == case["score"], len(case["seq"])
        assert [list(t) for t in tr] == case["trace"], len(case["seq"])
    seqs = [gen.rna_gc(n, 40 + n) for n in (0, 1, 5, 9, 31, 64, 150, 400)]
    got = cg.parse_batch(seqs)
    for s, r in zip(seqs, got):
        score, btr = ctx.zuker(s)
        assert r[0] == score and r[2] == btr, len(s)
    # an input past the tile's 8 KB copy: the windowed input reads of a
    # grammar that fills the LDS
    s = gen.rna_gc(8300, 77)
    print("zuker grammar n=%d (windowed input)" % len(s), flush=True)
    ans, found, tr = cg.parse(s)
    score, btr = ctx.zuker(s)
    assert found and ans == score and tr == btr


def test_rnafold_grammar_vs_builtin(ctx):
    """RNAFold.scala through the generic path equals dpx_rnafold (rule ids
    Q 0-3, QM 4, Qp 5-8) and the committed oracle fixture."""
    import dpx.examples as ex
    from dpx import gen
    g = ex.rnafold()
    cg = ctx.grammar(g)
    assert (cg.info.n_rules, cg.info.n_tabs, cg.info.rna) == (9, 3, 1)
    seqs = [gen.rna_gc(n, 7 * n + 3) for n in (0, 1, 2, 5, 12, 33, 64, 150, 300)] + ["AAAAAAAAAA", "GGGAAACCC"]
    got = cg.parse_batch(seqs)
    for s, r in zip(seqs, got):
        print("rnafold grammar n=%d" % len(s), flush=True)
        score, btr = ctx.rnafold(s)
        assert r == cg.parse(s)
        if r[1]:
            assert r[0] == score and r[2] == btr, len(s)
        else:  # no structure: the built-in reports 999999 and no trace
            assert (score, btr) == (999999, []), len(s)
    for fx in json.load(open(o.ROOT + "/tests/golden/rnafold_large.json")):  # oracle answers at n = 1500, 2048
        seq = getattr(gen, fx["gen"])(fx["n"], fx["seed"])
        print("rnafold grammar fixture n=%d" % len(seq), flush=True)
        ans, found, tr = cg.parse(seq)
        assert found and ans == fx["score"] and [list(t) for t in tr] == fx["trace"]



def test_repeated_shapes(ctx):
    """Different inputs of one shape through one compiled grammar, one after
    the other (the tables and instance records are reused), each equal to
    the built-in kernel."""
    from dpx import gen
    sw = ctx.grammar(G.swat_affine())
    for seed in range(4):
        a, b = gen.dna_pair(300, 280, 500 + seed)
        ans, found, tr = sw.parse(a, b)
        s, _, btr = ctx.swat(a, b, 2, -3, -5, -2)
        assert found and ans == s and [t[:3] for t in tr] == [t[:3] for t in btr], seed
    nu = ctx.grammar(G.nussinov())
    for seed in range(3):
        s = gen.rna_uniform(400, 900 + seed)
        assert nu.parse(s)[::2] == ctx.nussinov(s, min_loop=0), seed


def test_sum_count_closed_forms(ctx):
    """h = sum / count: Motzkin numbers (int64 and int32 answers, candidates
    split over lanes; doubles in sequential order on one lane), split counts,
    windows; no backtrace."""
    cases = [("long long", "q", 64, [0, 1, 7, 33, 200, 1100]), ("int", "i", 32, [2, 40, 300]),
             ("double", "d", None, [0, 20, 33])]
    for answer, fmt, bits, ns in cases:
        cg = ctx.grammar(G.motzkin(answer, fmt))
        want = G.motzkin_numbers(max(ns), bits)
        for n in ns:
            print("motzkin %s n=%d" % (answer, n), flush=True)
            assert bits or want[n] < 2 ** 53
            ans, found, tr = cg.parse("a" * n)
            assert found and ans == want[n] and tr == [], (answer, n)
        got = cg.parse_batch(["a" * n for n in ns])
        assert [x[0] for x in got] == [want[n] for n in ns]
    valid, count = G.split_counts(400)
    cg = ctx.grammar(G.splits())
    for n in [0, 1, 2, 3, 4, 5, 64, 65, 199, 400]:
        ans, found, tr = cg.parse("a" * n)
        assert found == valid[n] and (not found or ans == count[n]) and tr == [], n
    for w in [3, 4, 7]:
        cg = ctx.grammar(G.splits(window=w))
        for n in [2, 7, 50]:
            ans, found, _ = cg.parse("a" * n)
            assert found == (n >= w and valid[w]) and (not found or ans == n - w + 1), (w, n)
    mz = G.motzkin_numbers(9)
    cg = ctx.grammar(G.motzkin(window=9))
    for n in [5, 9, 100]:
        ans, found, _ = cg.parse("a" * n)
        assert found == (n >= 9) and (not found or ans == (n - 8) * mz[9]), n


def test_random_sum_count_grammars(ctx):
    """Random grammars under h = sum (product algebras) and h = count against
    the oracle: answers and found flags, batched and one by one."""
    import adp_random as R
    done = 0
    for seed in range(24):
        agg = "sum" if seed % 2 else "count"
        g = R.single_track(seed, agg) if seed % 3 else R.two_track(seed, agg)
        try:
            G.oracle(g)
        except ValueError:
            continue
        print("random %s seed %d" % (agg, seed), flush=True)
        cg = ctx.grammar(g)
        rng = random.Random(seed)
        ins = []
        for _ in range(4):
            if g.tracks == 1:
                ins.append((R.text(rng, rng.randint(0, 14)), None))
            else:
                ins.append((R.text(rng, rng.randint(0, 9)), R.text(rng, rng.randint(0, 9))))
        want = [G.run_oracle(g, a, b)[1] for a, b in ins]
        for (a, b), (oans, ofound, _) in zip(ins, want):
            ans, found, tr = cg.parse(a, b)
            assert found == bool(ofound) and tr == [], (seed, a, b)
            if found:
                assert ans == oans, (seed, a, b, ans, oans)
        got = cg.parse_batch([a if b is None else (a, b) for a, b in ins])
        for (ans, found, _), (oans, ofound, _) in zip(got, want):
            assert found == bool(ofound) and (not found or ans == oans), seed
        done += 1
    assert done >= 12


def test_zuker_count(ctx):
    """ZukerCount (Zuker.scala:28-44, h = sum) on the device against the
    oracle (itself checked against brute-force derivation counts), and a
    longer sequence whose count wraps in int32 both ways."""
    from dpx import gen
    g = G.zuker("count")
    cg = ctx.grammar(g)
    seqs = [gen.rna_uniform(n, 77 + n) for n in [0, 4, 7, 12, 20, 33, 45, 60]]
    for s in seqs:
        print("zuker count n=%d" % len(s), flush=True)
        _, (oans, ofound, _) = G.run_oracle(g, s)
        ans, found, tr = cg.parse(s)
        assert found and ofound and ans == oans and tr == [], len(s)
    got = cg.parse_batch(seqs)
    assert [x[0] for x in got] == [cg.parse(s)[0] for s in seqs]


def test_band_pipeline_edges(ctx):
    """Two-track grammars without split loops run as one band-pipeline launch
    (64 rows per band, columns published every 16 steps): shapes on every
    side of a band and a chunk boundary, singly and as one batch of mixed
    sizes, equal the built-in SW kernel (score and trace)."""
    from dpx import gen
    sc = (2, -3, -5, -2)
    cg = ctx.grammar(G.swat_affine(*sc))
    pairs = []
    for n1 in [0, 1, 63, 64, 65, 127, 128, 200]:
        for n2 in [0, 1, 15, 16, 17, 100]:
            pairs.append(gen.dna_pair(n1, n2, 1000 + 7 * n1 + n2))
    want = [ctx.swat(a, b, *sc) for a, b in pairs]
    for (a, b), (s, _, btr) in zip(pairs, want):
        ans, found, tr = cg.parse(a, b)
        assert found and ans == s and [t[:3] for t in tr] == [t[:3] for t in btr], (len(a), len(b))
    got = cg.parse_batch(pairs)
    for (a, b), (s, _, btr), (ans, found, tr) in zip(pairs, want, got):
        assert found and ans == s and [t[:3] for t in tr] == [t[:3] for t in btr], (len(a), len(b))


def test_band_pipeline_empty_value(ctx):
    """A two-track grammar with the reference's empty value and no
    always-valid table (validity read as `!= empty` from the band's
    registers): edit distance against the oracle (score and trace) and
    against the textbook recurrence at band and chunk edges, single and
    batched."""
    import adp_random as R
    g = G.edit_distance()
    cg = ctx.grammar(g)
    assert "struct Regs" in cg.source
    rng = random.Random(77)
    for n1, n2 in [(0, 0), (0, 5), (7, 0), (12, 9), (30, 25)]:
        _vs_oracle(cg, g, R.text(rng, n1, "ab"), R.text(rng, n2, "ab"))
    pairs = [(R.text(rng, n1, "acgt"), R.text(rng, n2, "acgt"))
             for n1, n2 in [(63, 200), (64, 17), (65, 100), (200, 150), (300, 64), (129, 16)]]
    for a, b in pairs:
        ans, found, _ = cg.parse(a, b)
        assert found and ans == G.levenshtein(a, b), (len(a), len(b))
    got = cg.parse_batch(pairs)
    assert [x[0] for x in got] == [G.levenshtein(a, b) for a, b in pairs]


def test_nested_aggregations_vs_oracle(ctx):
    """Nested aggregations (an aggregate below a concatenation, with the
    union semantics of the Scala parsers): seeded random grammars and a
    Nussinov variant with an aggregated pair interior, single parses (answer,
    found flag, full trace) and one batch against the oracle."""
    import adp_random as R
    import dpx
    from dpx import gen
    done = 0
    cases = [(R.nested_track(seed), seed) for seed in range(24)] + [(G.nested_pairs(), 500)]
    for g, seed in cases:
        try:
            G.oracle(g)
        except ValueError:
            with pytest.raises(dpx.DpxError):
                ctx.grammar(g)
            continue
        print("nested aggregation grammar seed %d" % seed, flush=True)
        cg = ctx.grammar(g)
        rng = random.Random(2000 + seed)
        ins = []
        for _ in range(4):
            s = R.text(rng, rng.randint(0, 20)) if seed < 500 else gen.rna_uniform(rng.randint(0, 60), rng.randint(0, 99))
            _vs_oracle(cg, g, s)
            ins.append(s)
        got = cg.parse_batch(ins)
        for q, s in enumerate(ins):
            assert got[q] == cg.parse(s), (seed, q)
        cg.close()
        done += 1
    assert done > 15


@pytest.mark.parametrize("wide", [False, True])
def test_middle_product_left_terms(ctx, wide):
    """A left term of the middle product (a call of the staged left operand
    alone), kept in LDS when its type is the answer type (int) and evaluated
    in place when it is not (long long): both against the oracle (answer and
    full trace) at sizes whose tiles have a middle product."""
    from dpx import gen
    g = G.left_terms(wide)
    cg = ctx.grammar(g)
    rng = random.Random(31 + int(wide))
    for n in (0, 1, 33, 97, 130, 200):
        s = gen.rna_uniform(n, rng.randint(0, 99))
        print("left-term grammar wide=%s n=%d" % (wide, n), flush=True)
        _vs_oracle(cg, g, s)
    cg.close()
