"""Backtrace replay through the reference's pretty-print algebras
(dpx.pretty), checked on the reference's golden traces and on oracle traces
with properties the algebras fix independently."""
import json
import os
import random

import pytest

import oracle_lib as o
from dpx import gen, pretty

HERE = os.path.dirname(os.path.abspath(__file__))
PAIRS = {("A", "U"), ("U", "A"), ("G", "C"), ("C", "G"), ("G", "U"), ("U", "G")}


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def _pairs(db):
    st, out = [], []
    for q, ch in enumerate(db):
        if ch == "(":
            st.append(q)
        elif ch == ")":
            out.append((st.pop(), q))
    assert not st
    return sorted(out)


def _marked_pairs(trace, zuker):
    """Pairs read off the records directly: a cl record (rule 0-5) closes
    (i, j-1), rules 1-5 also (i+1, j-2); a Nussinov pair record closes (i, j-1)."""
    out = set()
    for i, j, r, _, _ in trace:
        if zuker and 0 <= r <= 5:
            out.add((i, j - 1))
            if r >= 1:
                out.add((i + 1, j - 2))
        if not zuker and r == 3:
            out.add((i, j - 1))
    return sorted(out)


def _zt(tr):
    return [tuple(x) for x in tr]


@pytest.mark.parametrize("fx", _load("zuker_ref.json"), ids=lambda f: "n%d" % len(f["seq"]))
def test_zuker_reference_traces_build(fx):
    seq, tr = fx["seq"].upper(), _zt(fx["trace"])
    db = pretty.zuker(seq, tr)
    assert len(db) == len(seq)
    ps = _pairs(db)
    assert ps == _marked_pairs(tr, True)
    for i, j in ps:
        assert (seq[i], seq[j]) in PAIRS
        # the grammar has no lonely pairs: every pair is stacked on a neighbour
        assert (i + 1, j - 1) in ps or (i - 1, j + 1) in ps


@pytest.mark.parametrize("fx", _load("nussinov_ref.json"), ids=lambda f: "n%d" % len(f["seq"]))
def test_nussinov_reference_traces_build(fx):
    seq, tr = fx["seq"].upper(), _zt(fx["trace"])
    db = pretty.nussinov(seq, tr)
    assert len(db) == len(seq)
    ps = _pairs(db)
    assert len(ps) == fx["score"]  # each pair scores +1 (Nussinov.scala:49-56)
    assert ps == _marked_pairs(tr, False)
    for i, j in ps:
        assert (seq[i], seq[j]) in PAIRS and i + 1 <= j - 1 + 1


def test_matchain_reference_string():
    # MatrixMult.scala:131 / Tests.scala:24
    rows, cols = [10, 100, 5], [100, 5, 50]
    _, tr = o.matchain(rows, cols)
    assert pretty.matchain(rows, cols, tr) == "((|10x100|*|100x5|)*|5x50|)"


def test_matchain_string_cost():
    rnd = random.Random(4)
    for _ in range(20):
        n = rnd.randint(1, 12)
        rows, cols = gen.matrix_chain(n, rnd.randint(0, 10 ** 6), 1, 30)
        cost, tr = o.matchain(rows, cols)
        s = pretty.matchain(rows, cols, tr)

        def parse(p):  # evaluate the parenthesisation: (cost, rows, cols, next position)
            if s[p] == "|":
                q = s.index("|", p + 1)
                r, c = map(int, s[p + 1:q].split("x"))
                return 0, r, c, q + 1
            c1, r1, k1, p = parse(p + 1)
            assert s[p] == "*"
            c2, r2, k2, p = parse(p + 1)
            assert s[p] == ")" and k1 == r2
            return c1 + c2 + r1 * k1 * k2, r1, k2, p + 1

        c, _, _, end = parse(0)
        assert end == len(s) and c == cost


@pytest.mark.parametrize("sc", [(10, -3, -3, -1), (2, -3, -5, -2)])
def test_swat_tracks(sc):
    match, mismatch, go, ge = sc
    rnd = random.Random(11)
    for _ in range(30):
        a = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 40)))
        b = "".join(rnd.choice("ACGT") for _ in range(rnd.randint(1, 40)))
        for fb in (False, True):
            axiom, (best, bi, bj), tr = o.swat(a, b, *sc, from_best=fb)
            t1, t2 = pretty.swat(a, b, tr)
            # one column per pair / open / extend record, appended at the end;
            # the start or zero record in front gives "." and, for a zero
            # rule, '-' times the length of the other prefix (SWatAffine.scala:37,45-46)
            ncol = sum(1 for r in tr if r[2] in (1, 5, 6, 8, 9))
            s1, s2 = t1[len(t1) - ncol:], t2[len(t2) - ncol:]
            h1, h2 = t1[:len(t1) - ncol], t2[:len(t2) - ncol]
            assert h1[0] == "." and h2[0] == "." and set(h1[1:] + h2[1:]) <= {"-"}
            i0, j0 = tr[0][0], tr[0][1]
            assert s1.replace("-", "") == a[i0 - len(s1.replace("-", "")):i0]
            assert s2.replace("-", "") == b[j0 - len(s2.replace("-", "")):j0]
            # re-score the alignment column by column with Gotoh's affine gaps
            # and the reference's clamps at 0 (open/extend, SWatAffine.scala:26-29)
            score, g1, g2 = 0, None, None
            for x, y in zip(s1, s2):
                if x != "-" and y != "-":
                    score, g1, g2 = score + (match if x == y else mismatch), None, None
                elif y == "-":
                    g1 = max(0, (score if g1 is None else g1) + (go if g1 is None else ge))
                    score, g2 = g1, None
                else:
                    g2 = max(0, (score if g2 is None else g2) + (go if g2 is None else ge))
                    score, g1 = g2, None
            assert score == (best if fb else axiom), (a, b, t1, t2)


@pytest.mark.parametrize("seed", range(4))
def test_rnafold_build(seed):
    """Dot-bracket of RNAfold oracle traces: right length, canonical pairs,
    exactly the pairs the trace's Qp records close (rules 5-8 close (i, j-1)),
    and every unpaired stretch of an interior loop under 30 bases."""
    seq = gen.rna_gc(60 + 70 * seed, 400 + seed)
    score, tr = o.rnafold(seq)
    db = pretty.rnafold(seq, tr)
    assert len(db) == len(seq)
    ps = _pairs(db)
    assert ps == sorted({(i, j - 1) for i, j, r, _, _ in tr if 5 <= r <= 8})
    for i, j in ps:
        assert (seq[i], seq[j]) in PAIRS
    for i, j, r, k, l in tr:
        if r == 7:
            assert 0 <= k - 1 - i < 30 and 0 <= j - 1 - l < 30
    assert pretty.rnafold("AAAA", o.rnafold("AAAA")[1]) is None
