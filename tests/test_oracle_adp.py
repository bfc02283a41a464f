"""The generic-grammar oracle (oracle/adp_oracle.py) pinned on the CPU:

* against the reference's own generated code: the Nussinov grammar of
  Nussinov.scala must reproduce bench2/nu.c's golden vectors (score and full
  trace: rule ids, split positions, BFS order, ties);
* against the C oracle's Smith-Waterman/Gotoh and matrix chain (themselves
  pinned by brute force and the reference's values) for SWatAffine.scala and
  MatrixMult2.scala written as grammars;
* against brute force: the optimum over every derivation (the reference's
  Scala `apply` without aggregation) for ElMamun, Brackets (window), SeqAlign
  (both algebras), Nussinov, and every trace replayed (`build`) gives the
  answer;
* the library generates and compiles (hipRTC, no device) each grammar, and
  rejects what it does not support with a message.
"""
import ctypes
import json
import random

import pytest

import adp_grammars as G
import oracle_lib as o


def _ok_trace(orc, g, in1, in2, ans, tr):
    assert orc.replay(g.py_input(in1), g.py_input(in2) if in2 is not None else None, tr) == ans


def test_nussinov_grammar_matches_reference_generated_code():
    fx = json.load(open(o.ROOT + "/tests/golden/nussinov_ref.json"))
    g = G.nussinov()
    for case in fx:
        if len(case["seq"]) > 60:
            continue
        orc, (ans, found, tr) = G.run_oracle(g, case["seq"])
        assert found and ans == case["score"], case["seq"]
        assert [list(t) for t in tr] == case["trace"], case["seq"]


@pytest.mark.parametrize("n1,n2", [(0, 0), (0, 3), (4, 0), (1, 1), (5, 9), (12, 7), (17, 17)])
def test_swat_affine_grammar_matches_c_oracle(n1, n2):
    from dpx import gen
    a, b = gen.dna_pair(n1, n2, 100 * n1 + n2)
    for sc in [(2, -3, -5, -2), (10, -3, -3, -1)]:
        g = G.swat_affine(*sc)
        orc, (ans, found, tr) = G.run_oracle(g, a, b)
        s, best, otr = o.swat(a, b, *sc, from_best=False)
        assert found and ans == s
        assert [t[:3] for t in tr] == [t[:3] for t in otr]
        _ok_trace(orc, g, a, b, ans, tr)


def test_matrix_chain_grammar_matches_c_oracle():
    rng = random.Random(5)
    for n in [1, 2, 3, 5, 9, 14]:
        dims = [rng.randint(1, 60) for _ in range(n + 1)]
        mats = [(dims[q], dims[q + 1]) for q in range(n)]
        g = G.matrix_chain()
        orc, (ans, found, tr) = G.run_oracle(g, mats)
        s, otr = o.matchain([m[0] for m in mats], [m[1] for m in mats])
        assert found and ans == s
        assert [t[:4] for t in tr] == [t[:4] for t in otr]
        _ok_trace(orc, g, mats, None, ans, tr)
    # the reference's test values (Tests.scala:16-25)
    for mats, want in [([(10, 100), (100, 5), (5, 50)], 7500), ([(1, 2), (2, 20), (20, 2), (2, 4), (4, 2), (2, 1), (1, 7), (7, 3)], 122)]:
        _, (ans, _, _) = G.run_oracle(G.matrix_chain(), mats)
        assert ans == want


def _brute(g, in1, in2=None):
    orc, (ans, found, tr) = G.run_oracle(g, in1, in2)
    vals = orc.enumerate_all(g.py_input(in1), g.py_input(in2) if in2 is not None else None)
    return orc, ans, found, tr, vals


@pytest.mark.parametrize("expr", ["1+2*3*4+5", "9", "2*3", "1+1*1+1*2", "3*0+4*2*1", "+", "1+"])
def test_elmamun_vs_brute_force(expr):
    for h, pick in [("min", min), ("max", max)]:
        g = G.elmamun(h)
        orc, ans, found, tr, vals = _brute(g, expr)
        assert found == bool(vals)
        if vals:
            assert ans == pick(vals)
            _ok_trace(orc, g, expr, None, ans, tr)
    if expr == "1+2*3*4+5":  # ADP's classic: the seller gets (1+2)*3*(4+5), the buyer 1+(2*3*4)+5
        assert G.run_oracle(G.elmamun("max"), expr)[1][0] == 81
        assert G.run_oracle(G.elmamun("min"), expr)[1][0] == 30


def test_brackets_window_vs_brute_force():
    rng = random.Random(3)
    cases = ["(((3)))((2))(1)((6))((((8))))(3)((4))", "((1)(2))", "(5)", "1234", "((((((((9))))))))"]
    for _ in range(8):
        cases.append("".join(rng.choice("()123") for _ in range(rng.randint(3, 14))))
    for s in cases:
        g = G.brackets(8)
        orc, ans, found, tr, vals = _brute(g, s)
        assert found == bool(vals), s
        if vals:
            assert ans == max(vals), s
            _ok_trace(orc, g, s, None, ans, tr)


@pytest.mark.parametrize("alg", ["sw", "nw"])
def test_seqalign_vs_brute_force(alg):
    from dpx import gen
    for n1, n2, seed in [(0, 0, 1), (1, 2, 2), (3, 3, 3), (4, 2, 4), (5, 4, 5)]:
        a, b = gen.dna_pair(n1, n2, seed)
        g = G.seqalign(alg)
        orc, ans, found, tr, vals = _brute(g, a, b)
        assert found and ans == max(vals)
        _ok_trace(orc, g, a, b, ans, tr)


def test_nussinov_vs_brute_force():
    rng = random.Random(9)
    for n in list(range(0, 7)) * 2:  # the grammar is ambiguous: derivations grow exponentially
        s = "".join(rng.choice("ACGU") for _ in range(n))
        g = G.nussinov()
        orc, ans, found, tr, vals = _brute(g, s)
        assert ans == max(vals)
        _ok_trace(orc, g, s, None, ans, tr)


def _check(spec):
    L = ctypes.CDLL(o.ROOT + "/lamp-dp-mt_amd/libdpx.so")
    L.dpx_grammar_check.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_char_p), ctypes.POINTER(ctypes.c_char_p)]
    src, msg = ctypes.c_char_p(), ctypes.c_char_p()
    rc = L.dpx_grammar_check(spec.encode(), ctypes.byref(src), ctypes.byref(msg))
    return rc, (src.value or b"").decode(), (msg.value or b"").decode()


@pytest.mark.parametrize("make", [G.nussinov, lambda: G.elmamun("min"), G.brackets, G.seqalign,
                                  lambda: G.seqalign("nw"), G.swat_affine, G.matrix_chain])
def test_library_generates_and_compiles(make):
    g = make()
    rc, src, msg = _check(g.spec_json())
    assert rc == 0, msg
    # one track: the tile queue; two tracks: diagonal launches (and the band
    # pipeline where there are no split loops)
    assert ("adp_tile_1024" in src) == (g.tracks == 1) and ("adp_fill_1" in src) == (g.tracks == 2)
    assert "adp_finish" in src


def test_library_rejects_unsupported():
    g = G.nussinov()
    spec = g.spec()
    bad = dict(spec, h="product")
    rc, _, msg = _check(json.dumps(bad))
    assert rc != 0 and "aggregation" in msg
    bad = dict(spec, axiom="nope")
    assert _check(json.dumps(bad))[0] != 0
    rc, _, msg = _check("{not json")
    assert rc != 0 and "JSON" in msg
    rc, _, msg = _check("[" * 100000)  # nesting is bounded, not a stack overflow
    assert rc != 0 and "deeply" in msg
    # a body that does not compile: the compiler log comes back
    bad = json.loads(g.spec_json())
    fn = sorted(bad["functions"])[0]
    bad["functions"][fn]["body"] = "return undeclared_name;"
    rc, _, msg = _check(json.dumps(bad))
    assert rc != 0 and "undeclared_name" in msg
    # a cycle within one cell
    cyc = {"tracks": 1, "h": "max", "axiom": "A", "functions": {},
           "tabulations": {"A": {"p": {"tab": "B"}}, "B": {"p": {"tab": "A"}}}}
    rc, _, msg = _check(json.dumps(cyc))
    assert rc != 0 and "loop" in msg


def test_random_grammars_oracle_vs_brute_force():
    """Seeded random grammars (every construct) against brute force and
    trace replay; grammars the analysis rejects (same-cell cycles) skipped."""
    import adp_random as R
    checked = 0
    for seed in range(60):
        g = R.single_track(seed) if seed % 3 else R.two_track(seed)
        try:
            orc = G.oracle(g)
        except ValueError:
            continue
        if g.tracks == 2 and any(t.always_valid for t in g.tabs.values()):
            continue  # an always_valid table reads 0 from cells without a derivation: no brute-force meaning
        if isinstance(g.h, tuple):
            continue  # maxBy a non-monotone key: the DP answer is not the best derivation (checked on the GPU only)
        rng = random.Random(seed)
        for _ in range(3):
            if g.tracks == 1:
                in1, in2 = R.text(rng, rng.randint(0, 6)), None
            else:
                in1, in2 = R.text(rng, rng.randint(0, 3)), R.text(rng, rng.randint(0, 3))
            _, (ans, found, tr) = G.run_oracle(g, in1, in2)
            try:
                vals = orc.enumerate_all(g.py_input(in1), g.py_input(in2) if in2 is not None else None, limit=50000)
            except OverflowError:
                continue
            if g.empty_value is not None:
                ev = int(g.empty_value)
                vals = [v for v in vals if v != ev]
            assert found == bool(vals), (seed, in1, in2)
            if vals:
                key = orc._key
                want = max(vals, key=key) if (g.h in ("max",) or isinstance(g.h, tuple)) else min(vals, key=key)
                assert key(ans) == key(want), (seed, in1, in2, ans, want)
                assert key(orc.replay(g.py_input(in1), g.py_input(in2) if in2 is not None else None, tr)) == key(ans)
            checked += 1
    assert checked > 60


def test_random_grammars_compile():
    import adp_random as R
    for seed in range(0, 60, 7):
        g = R.single_track(seed) if seed % 3 else R.two_track(seed)
        rc, src, msg = _check(g.spec_json())
        assert rc == 0 or "loop" in msg, (seed, msg[:2000])


def test_random_grammars_compile_all():
    """Every random grammar the oracle accepts generates and compiles."""
    import adp_random as R
    for seed in range(48):
        g = R.single_track(seed) if seed % 3 else R.two_track(seed)
        try:
            G.oracle(g)
        except ValueError:
            continue
        rc, src, msg = _check(g.spec_json())
        assert rc == 0, (seed, msg[:2000])


def test_sum_aggregation_closed_form():
    """h = sum with the counting algebra: Motzkin numbers (exact, then int32
    and int64 wrapping as the answer type does), windows add up."""
    exact = G.motzkin_numbers(40)
    assert exact[:11] == [1, 1, 2, 4, 9, 21, 51, 127, 323, 835, 2188]  # OEIS A001006
    for answer, fmt, bits in [("int", "i", 32), ("long long", "q", 64)]:
        g = G.motzkin(answer, fmt)
        want = G.motzkin_numbers(40, bits)
        for n in [0, 1, 2, 5, 11, 25, 40]:
            _, (ans, found, tr) = G.run_oracle(g, "a" * n)
            assert found and ans == want[n] and tr == [], (answer, n)
    g = G.motzkin(h="sum", window=6)
    for n in [3, 6, 13]:
        _, (ans, found, _) = G.run_oracle(g, "a" * n)
        assert found == (n >= 6) and (not found or ans == (n - 5) * exact[6])


def test_count_aggregation_closed_form():
    valid, count = G.split_counts(30)
    g = G.splits()
    for n in range(0, 31, 3):
        _, (ans, found, tr) = G.run_oracle(g, "a" * (n or 1) if n == 1 else "a" * n)
        assert found == valid[n] and (not found or ans == count[n]) and tr == [], n
    for w in [3, 4, 7]:
        g = G.splits(window=w)
        for n in [2, 7, 12]:
            _, (ans, found, _) = G.run_oracle(g, "a" * n)
            assert found == (n >= w and valid[w]) and (not found or ans == n - w + 1), (w, n)
    # sum over the same grammar adds the candidates' values
    _, (ans, found, _) = G.run_oracle(G.splits(h="sum"), "a" * 3)
    assert found and ans == 2  # el (s=1) (s=1) -> 1 + 1


def test_random_sum_grammars_vs_brute_force():
    """Product algebras under h = sum: the DP answer is the sum of every
    derivation's value, modulo 2^32 (distributivity)."""
    import adp_random as R
    checked = 0
    for seed in range(60):
        g = R.single_track(seed, "sum") if seed % 3 else R.two_track(seed, "sum")
        try:
            orc = G.oracle(g)
        except ValueError:
            continue
        av = any(t.always_valid for t in g.tabs.values())
        rng = random.Random(seed)
        for _ in range(3):
            if g.tracks == 1:
                in1, in2 = R.text(rng, rng.randint(0, 6)), None
            else:
                in1, in2 = R.text(rng, rng.randint(0, 3)), R.text(rng, rng.randint(0, 3))
            _, (ans, found, tr) = G.run_oracle(g, in1, in2)
            try:
                vals = orc.enumerate_all(g.py_input(in1), g.py_input(in2) if in2 is not None else None, limit=50000)
            except OverflowError:
                continue
            assert tr == []
            if not av:
                assert found == bool(vals), (seed, in1, in2)
            if found:
                assert ans == R._wrap(sum(vals)), (seed, in1, in2, ans, vals)
            checked += 1
    assert checked > 60


def test_zuker_count_vs_brute_force():
    """ZukerCount (h = sum): the number of derivations of the Zuker grammar."""
    from dpx import gen
    g = G.zuker("count")
    for n in [0, 5, 9, 12, 14]:
        for seed in range(3):
            s = gen.rna_uniform(n, 31 * n + seed)
            orc, (ans, found, tr) = G.run_oracle(g, s)
            vals = orc.enumerate_all(g.py_input(s), limit=400000)
            assert found and ans == len(vals) and all(v == 1 for v in vals), (n, seed)
    # structures appear: a sequence that can fold has more than the open chain
    s = "GGGGAAAACCCC"
    orc, (ans, _, _) = G.run_oracle(g, s)
    assert ans > 1


def test_sum_count_grammars_compile():
    import adp_random as R
    gs = [G.motzkin(), G.motzkin("double", "d"), G.splits(), G.splits(window=4), G.zuker("count")]
    for seed in range(12):
        g = R.single_track(seed, "count") if seed % 3 else R.two_track(seed, "sum")
        try:
            G.oracle(g)
        except ValueError:
            continue
        gs.append(g)
    assert len(gs) > 10
    for g in gs:
        rc, src, msg = _check(g.spec_json())
        assert rc == 0, msg[:2000]


def test_fill_strategy_of_generated_code():
    """Two-track grammars without split loops get the band-pipeline kernel,
    in its register form when every reference is (i, j - c) or (i - 1, j - c);
    two-track grammars with split loops keep one launch per diagonal; one-track
    grammars run on the tile queue alone."""
    import dpx
    for g, band, regs in [(G.swat_affine(), True, True), (G.seqalign("sw"), False, False),
                          (G.nussinov(), False, False), (G.matrix_chain(), False, False)]:
        rc, src, msg = _check(g.spec_json())
        assert rc == 0, msg
        assert ("void adp_band(" in src) == band and ("struct Regs" in src) == regs, g
        assert ("adp_fill_1" in src) == (g.tracks == 2) and ("adp_tile_1024" in src) == (g.tracks == 1)
        assert "adp_finish" in src


def test_edit_distance_grammar_oracle_vs_textbook():
    """The two-track edit-distance grammar with an empty value: the oracle's
    answer is the Levenshtein distance, and its trace replays to it."""
    rng = random.Random(5)
    g = G.edit_distance()
    for _ in range(40):
        a = "".join(rng.choice("ab") for _ in range(rng.randint(0, 9)))
        b = "".join(rng.choice("ab") for _ in range(rng.randint(0, 9)))
        orc, (ans, found, tr) = G.run_oracle(g, a, b)
        assert found and ans == G.levenshtein(a, b), (a, b)
        assert orc.replay(g.py_input(a), g.py_input(b), tr) == ans


def test_nested_aggregation_vs_brute_force():
    """Nested aggregations (CodeGen.scala:31-36,108-135): the oracle against
    every derivation under the nested aggregations' own semantics (the first
    best of each nested list, BaseParsers.scala:228), and its trace replays
    to its answer; random grammars with monotone algebras (Bellman holds) and
    a Nussinov variant with an aggregated pair interior."""
    import adp_random as R
    checked = 0
    cases = [(R.nested_track(seed), seed) for seed in range(40)] + [(G.nested_pairs(), 100)]
    for g, seed in cases:
        try:
            orc = G.oracle(g)
        except ValueError:
            continue
        rng = random.Random(seed)
        for _ in range(3):
            in1 = R.text(rng, rng.randint(0, 7)) if seed < 100 else "".join(rng.choice("ACGU") for _ in range(rng.randint(0, 9)))
            _, (ans, found, tr) = G.run_oracle(g, in1)
            try:
                vals = orc.enumerate_all(g.py_input(in1), limit=50000, aggregate_tabs=True)
            except OverflowError:
                continue
            assert found == bool(vals), (seed, in1)
            if vals:
                want = max(vals) if g.h == "max" else min(vals)
                assert ans == want, (seed, in1, ans, want)
                assert orc.replay(g.py_input(in1), None, tr) == ans
            checked += 1
    assert checked > 60


def test_nested_aggregation_grammars_compile():
    import adp_random as R
    for seed in range(0, 40, 3):
        rc, src, msg = _check(R.nested_track(seed).spec_json())
        assert rc == 0, (seed, msg[:2000])
        assert "_g0c" in src  # a nested aggregation's block
    rc, src, msg = _check(G.nested_pairs().spec_json())
    assert rc == 0, msg[:2000]
    # nested sums are not supported (the reference's emitted code for them
    # adds into an uninitialised accumulator, CodeGen.scala:126-127)
    import dpx.adp as A
    g = A.Grammar(tracks=1, h="sum")
    one = A.cfun(lambda: 1, "", "return 1;")
    f = A.cfun(lambda c, x: x, "c,x", "return x;")
    s = g.tabulate("s", lambda: ((g.empty ^ one) | (g.el * (g.el * s ^ f).aggregate("sum") ^ f)).aggregate("sum"))
    g.axiom = s
    rc, _, msg = _check(g.spec_json())
    assert rc != 0 and "nested" in msg


def test_middle_product_left_terms_generated():
    """Zuker.scala's multiloop term of the closing pair (i, k) is generated as
    a left term of the middle product (one value per row and split, staged in
    LDS); a left term of another type than the answer compiles too (it is
    then evaluated in place)."""
    rc, src, msg = _check(G.zuker().spec_json())
    assert rc == 0, msg
    assert "mu0_<MUH0_>(" in src and "struct MidSm : MidSmB" in src
    for wide in (False, True):
        rc, src, msg = _check(G.left_terms(wide).spec_json())
        assert rc == 0, msg
        assert "mu0_<MUH0_>(" in src
