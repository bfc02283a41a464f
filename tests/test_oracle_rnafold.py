"""RNAfold grammar (RNAFold.scala:20-64) in the oracle: the MFE equals the
minimum, over every secondary structure the grammar derives, of the sum of
its loop energies (brute force on short sequences), and the backtrace spells
a structure of exactly that energy."""
import itertools
import random

import pytest

import oracle_lib as o

PAIRS = {("A", "U"), ("U", "A"), ("G", "C"), ("C", "G"), ("G", "U"), ("U", "G")}


def _structures(seq):
    """All non-crossing sets of canonical pairs (p, q) with q - p >= 2."""
    n = len(seq)
    memo = {}

    def rec(i, j):  # structures on [i, j)
        if (i, j) in memo:
            return memo[(i, j)]
        if j - i < 3:
            memo[(i, j)] = [()]
            return memo[(i, j)]
        out = list(rec(i + 1, j))  # i unpaired
        for q in range(i + 2, j):
            if (seq[i], seq[q]) in PAIRS:
                for a in rec(i + 1, q):
                    for b in rec(q + 1, j):
                        out.append(((i, q),) + a + b)
        memo[(i, j)] = out
        return out

    return rec(0, n)


def _energy(seq, pairs, fns):
    """Sum of loop energies under the grammar, None if it cannot derive the structure."""
    hl, il, _ = fns
    ps = sorted(pairs)
    e = 0
    for (p, q) in ps:
        inner = [(a, b) for (a, b) in ps if p < a and b < q and not any(p < c < a and b < d < q for (c, d) in ps)]
        if not inner:
            e += hl(p, q)
        elif len(inner) == 1:
            a, b = inner[0]
            if a - p - 1 > 29 or q - b - 1 > 29:
                return None
            e += il(p, a, b, q)
        # multiloops (2+ branches) cost nothing in this algebra
    return e


@pytest.mark.parametrize("seed", range(6))
def test_rnafold_mfe_brute_force(seed):
    rnd = random.Random(seed)
    for _ in range(12):
        n = rnd.randint(0, 13)
        seq = "".join(rnd.choice("ACGU") for _ in range(n))
        fns = o.energy_fns(seq)
        best = None
        for st in _structures(seq):
            if not st:
                continue
            e = _energy(seq, st, fns)
            if e is not None and (best is None or e < best):
                best = e
        score, tr = o.rnafold(seq)
        assert score == (o.RF_EMPTY if best is None else best), (seq, score, best)
        if best is not None:
            # the trace's closed pairs (rules 5-8 close (i, j-1)) form a structure of that energy
            pairs = sorted({(i, j - 1) for (i, j, r, _, _) in tr if 5 <= r <= 8})
            assert _energy(seq, pairs, fns) == score, (seq, pairs)


def test_rnafold_gc_rich_long_loops():
    # long hairpins / interior loops near the 30-base window edges
    seq = "GGGGG" + "A" * 31 + "CCCCC" + "AAAA" + "GGG" + "A" * 29 + "CCC"
    score, tr = o.rnafold(seq)
    fns = o.energy_fns(seq)
    pairs = sorted({(i, j - 1) for (i, j, r, _, _) in tr if 5 <= r <= 8})
    assert _energy(seq, pairs, fns) == score < 0


def test_rnafold_large_fixture_matches_generator_and_oracle_prefix():
    """The committed large fixtures name generator calls that still produce
    sequences the oracle folds to a trace starting at the axiom (0, n)."""
    import json
    import os
    from dpx import gen
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rnafold_large.json")) as f:
        fx = json.load(f)
    for c in fx:
        seq = getattr(gen, c["gen"])(c["n"], c["seed"])
        assert len(seq) == c["n"]
        assert c["trace"][0][:2] == [0, c["n"]] and c["score"] < 0
