"""Pins the parts of the oracle that no reference golden vector covers:
brute force on tiny inputs, closed forms and the reference-held values."""
import itertools
import random

import pytest

import oracle_lib as o

NEG = -10 ** 9


# ---------------------------------------------------------------- SW / Gotoh
def _best_alignment(a, b, match, mismatch, go, ge, must_end):
    """Brute force: every alignment (all column sequences) of every pair of
    substrings; returns the best score (0 = empty alignment).  A gap run of
    length L costs go + (L-1)*ge.  must_end: the substrings end at (len a, len b)."""
    best = 0
    n1, n2 = len(a), len(b)
    for p in range(n1 + 1):
        for q in range(p, n1 + 1):
            if must_end and q != n1:
                continue
            for r in range(n2 + 1):
                for s in range(r, n2 + 1):
                    if must_end and s != n2:
                        continue
                    x, y = a[p:q], b[r:s]

                    def rec(i, j, prev):
                        if i == len(x) and j == len(y):
                            return 0
                        out = NEG
                        if i < len(x) and j < len(y):
                            out = max(out, (match if x[i] == y[j] else mismatch) + rec(i + 1, j + 1, "M"))
                        if i < len(x):
                            out = max(out, (ge if prev == "A" else go) + rec(i + 1, j, "A"))
                        if j < len(y):
                            out = max(out, (ge if prev == "B" else go) + rec(i, j + 1, "B"))
                        return out

                    if x or y:
                        best = max(best, rec(0, 0, ""))
    return best


@pytest.mark.parametrize("scores", [(10, -3, -3, -1), (2, -3, -5, -2)])
def test_swat_local_score_brute_force(scores):
    rnd = random.Random(7)
    for _ in range(25):
        a = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(0, 4)))
        b = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(0, 4)))
        axiom, (best, bi, bj), _ = o.swat(a, b, *scores, trace=False)
        assert best == _best_alignment(a, b, *scores, must_end=False), (a, b)
        assert axiom == _best_alignment(a, b, *scores, must_end=True), (a, b)


@pytest.fixture(scope="module")
def sw_check(tmp_path_factory):
    """tests/sw_check.c built with gcc: explicit enumeration of alignments
    ("brute") and the textbook three-matrix Gotoh ("gotoh")."""
    import os
    import subprocess
    exe = str(tmp_path_factory.mktemp("swc") / "sw_check")
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sw_check.c")
    subprocess.check_call(["gcc", "-O2", "-std=c99", "-o", exe, src])

    def run(mode, cases):
        lines = "".join("%s %s %s %d %d %d %d\n" % (mode, a or "-", b or "-", *sc) for a, b, sc in cases)
        out = subprocess.run([exe], input=lines.encode(), capture_output=True, check=True).stdout.decode().split("\n")
        return [tuple(int(x) for x in line.split()) for line in out if line.strip()]

    return run


def _rand_scores(rnd):
    """Random scores under the affine-gap condition open <= ext <= 0: with a
    cheaper open than extension both the reference's recurrence and the
    textbook one would close and reopen a gap run, which the cost model of
    whole gap runs does not describe."""
    go = rnd.randint(-12, 0)
    return (rnd.randint(1, 10), rnd.randint(-10, 0), go, rnd.randint(go, 0))


def test_swat_brute_force_up_to_8(sw_check):
    """Every alignment of every substring pair enumerated (n <= 8): the
    oracle's max M is the best local score and M(n1,n2) the best local
    alignment ending at (n1,n2), for the two reference score sets and random ones."""
    rnd = random.Random(11)
    cases = []
    for k in range(60):
        n1, n2 = (8, 8) if k < 8 else (rnd.randint(0, 8), rnd.randint(0, 8))
        a = "".join(rnd.choice("ACGT") for _ in range(n1))
        b = "".join(rnd.choice("ACGT") for _ in range(n2))
        sc = [(10, -3, -3, -1), (2, -3, -5, -2)][k % 2] if k < 20 else _rand_scores(rnd)
        cases.append((a, b, sc))
    for (a, b, sc), (best, end) in zip(cases, sw_check("brute", cases)):
        axiom, (obest, _, _), _ = o.swat(a, b, *sc, trace=False)
        assert (obest, axiom) == (best, end), (a, b, sc)


def test_swat_equals_textbook_gotoh(sw_check):
    """max M equals the textbook local Gotoh score, and M(n1,n2) its H(n1,n2),
    at sizes up to 2000, random scores (SWatAffine.scala:22-35: the clamps of
    the gap tables at 0 do not change M)."""
    from dpx import gen
    rnd = random.Random(12)
    cases = []
    for k in range(24):
        n1, n2 = rnd.randint(0, 2000), rnd.randint(0, 2000)
        if k % 4 == 0:  # related pair: a long shared segment, high local scores
            base = gen.dna_uniform(n1 + 400, 90 + k)
            a, b = base[:n1], gen.dna_uniform(200, 190 + k) + base[n1 // 3: n1] + gen.dna_uniform(100, 290 + k)
        else:
            a, b = gen.dna_pair(n1, n2, 80 + k)
        sc = [(10, -3, -3, -1), (2, -3, -5, -2)][k % 2] if k < 4 else _rand_scores(rnd)
        cases.append((a, b, sc))
    for (a, b, sc), (best, end) in zip(cases, sw_check("gotoh", cases)):
        axiom, (obest, _, _), _ = o.swat(a, b, *sc, trace=False)
        assert (obest, axiom) == (best, end), (len(a), len(b), sc)


def test_swat_checker_catches_planted_mistakes(sw_check):
    """The two checkers disagree with a wrong score (a sign, a dropped gap
    extension), so they can tell such a mistake apart."""
    rnd = random.Random(13)
    cases = []
    for _ in range(20):  # b = a with 1-3 bases deleted in the middle: the best alignment has a gap run
        a = "".join(rnd.choice("ACGT") for _ in range(8))
        d = rnd.randint(1, 3)
        b = a[:3] + a[3 + d:]
        cases.append((a, b, (10, -3, -5, -2)))
    brute = sw_check("brute", cases)
    assert brute == sw_check("gotoh", cases)
    for a, b, _ in cases:
        assert o.swat(a, b, 10, -3, -5, -2, trace=False)[1][0] == brute[cases.index((a, b, _))][0]
    assert sw_check("gotoh", [(a, b, (10, -3, -5, -5)) for a, b, _ in cases]) != brute
    assert sw_check("gotoh", [(a, b, (10, -3, -2, -2)) for a, b, _ in cases]) != brute


def _trace_score(a, b, tr, match, mismatch, go, ge):
    """Score of the alignment a trace spells, by the SWatAffine pretty algebra:
    rule 1 pairs, rules 5/6 (8/9) open/extend a gap in track 1 (2), with the
    clamp at 0 of open/ext (SWatAffine.scala:26-29) applied bottom-up."""
    # walk from the deepest record back to the root, as the algebra would
    vals = {}
    for (i, j, rule, _, _) in reversed(tr):
        if rule in (0, 4, 7):
            v = 0
        elif rule == 1:
            v = vals[(i - 1, j - 1, "M")] + (match if a[i - 1] == b[j - 1] else mismatch)
        elif rule == 2:
            v = vals[(i, j, "g1")]
        elif rule == 3:
            v = vals[(i, j, "g2")]
        elif rule == 5:
            v = max(0, vals[(i - 1, j, "M")] + go)
        elif rule == 6:
            v = max(0, vals[(i - 1, j, "g1")] + ge)
        elif rule == 8:
            v = max(0, vals[(i, j - 1, "M")] + go)
        else:
            v = max(0, vals[(i, j - 1, "g2")] + ge)
        tab = "M" if rule <= 3 else "g1" if rule <= 6 else "g2"
        vals[(i, j, tab)] = v
    i, j, _, _, _ = tr[0]
    return vals[(i, j, "M")]


def test_swat_trace_replays_to_score():
    rnd = random.Random(3)
    for _ in range(40):
        a = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 30)))
        b = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 30)))
        for sc in [(10, -3, -3, -1), (2, -3, -5, -2)]:
            axiom, (best, bi, bj), tr = o.swat(a, b, *sc)
            assert tr[0][:2] == (len(a), len(b))
            assert _trace_score(a, b, tr, *sc) == axiom
            _, _, tr2 = o.swat(a, b, *sc, from_best=True)
            assert tr2[0][:2] == (bi, bj)
            assert _trace_score(a, b, tr2, *sc) == best


# -------------------------------------------------------------- matrix chain
def _chain_costs(rows, cols):
    """min over all parenthesisations (CLRS), int64, no overflow."""
    n = len(rows)

    def best(i, j):
        if j == i + 1:
            return 0
        return min(best(i, k) + best(k, j) + rows[i] * rows[k] * cols[j - 1] for k in range(i + 1, j))

    return best(0, n)


def test_matchain_reference_values():
    # Tests.scala:16-25: (10,7500,50) and (1,122,3)
    assert o.matchain([10, 100, 5], [100, 5, 50])[0] == 7500
    rows = [1, 2, 20, 2, 4, 2, 1, 7]
    cols = [2, 20, 2, 4, 2, 1, 7, 3]
    assert o.matchain(rows, cols)[0] == 122


def test_matchain_trace_builds_reference_string():
    # MatrixMult.scala:131 / Tests.scala:24: ((|10x100|*|100x5|)*|5x50|)
    rows, cols = [10, 100, 5], [100, 5, 50]
    _, tr = o.matchain(rows, cols)

    def build(q):
        i, j, rule, k, _ = tr[q]
        if rule == 0:
            return "|%dx%d|" % (rows[i], cols[i])
        kids = [p for p in range(len(tr)) if tr[p][:2] in ((i, k), (k, j)) and p > q]
        left = next(p for p in kids if tr[p][:2] == (i, k))
        right = next(p for p in kids if tr[p][:2] == (k, j))
        return "(" + build(left) + "*" + build(right) + ")"

    assert build(0) == "((|10x100|*|100x5|)*|5x50|)"


def test_matchain_brute_force():
    rnd = random.Random(5)
    for n in range(1, 8):
        for _ in range(5):
            dims = [rnd.randint(1, 30) for _ in range(n + 1)]
            rows, cols = dims[:-1], dims[1:]
            assert o.matchain(rows, cols)[0] == _chain_costs(rows, cols)


def test_matchain_int32_wrap():
    # three matrices whose split costs exceed 2**31: JVM Int wraps
    rows, cols = [2000, 60000, 30000], [60000, 30000, 50000]
    def wrap(v):
        return (v + 2 ** 31) % 2 ** 32 - 2 ** 31
    # M(0,2)=r0*r1*c1, M(1,3)=r1*r2*c2; candidates k=1,2 (MatrixMult2.scala:16)
    m02, m13 = wrap(2000 * 60000 * 30000), wrap(60000 * 30000 * 50000)
    c1 = wrap(0 + m13 + 2000 * 60000 * 50000)
    c2 = wrap(m02 + 0 + 2000 * 30000 * 50000)
    assert o.matchain(rows, cols)[0] == min(c1, c2)


# ------------------------------------------------------------------ Nussinov
def _nussinov_brute(seq, m):
    pairs = {("A", "U"), ("U", "A"), ("G", "U"), ("U", "G"), ("C", "G"), ("G", "C")}
    n = len(seq)
    cand = [(i, j) for i in range(n) for j in range(i + 1 + m, n) if (seq[i], seq[j]) in pairs]
    best = 0
    # every non-crossing, non-overlapping subset of candidate pairs
    for r in range(len(cand), 0, -1):
        if r <= best:
            break
        for sub in itertools.combinations(cand, r):
            used = set()
            ok = True
            for (i, j) in sub:
                if i in used or j in used:
                    ok = False
                    break
                used.add(i); used.add(j)
            if not ok:
                continue
            if all(not (i1 < i2 < j1 < j2 or i2 < i1 < j2 < j1) for (i1, j1), (i2, j2) in itertools.combinations(sub, 2)):
                best = r
                break
    return best


@pytest.mark.parametrize("m", [0, 1, 3])
def test_nussinov_brute_force(m):
    rnd = random.Random(11 + m)
    for _ in range(12):
        n = rnd.randint(0, 11)
        seq = "".join(rnd.choice("ACGU") for _ in range(n))
        assert o.nussinov(seq, min_loop=m, trace=False)[0] == _nussinov_brute(seq, m), seq
