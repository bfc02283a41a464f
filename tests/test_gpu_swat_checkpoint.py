"""Checkpointed Smith-Waterman traceback: with the traceback-code budget at
zero, dpx_swat keeps only checkpoint rows and recomputes one segment of
bands at a time during the walk; scores, best cells and traces must be the
oracle's, as with every code kept.  Also alignments taller than the
boundary-row ring (more than 4097 bands of 64 rows), where its slots are
reused."""
import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def _check(ctx, a, b, sc, tf):
    s, best, tr = ctx.swat(a, b, *sc, trace_from=tf)
    os_, obest, otr = o.swat(a, b, *sc, from_best=bool(tf))
    assert (s, best) == (os_, obest), (len(a), len(b), tf)
    assert [x[:3] for x in tr] == [x[:3] for x in otr], (len(a), len(b), tf)


@pytest.mark.parametrize("n1,n2", [(1, 5), (63, 64), (64, 200), (65, 129), (700, 333), (2000, 1500), (4097, 301)])
def test_checkpointed_trace_vs_oracle(ctx, n1, n2):
    from dpx import gen
    print("swat checkpointed %dx%d" % (n1, n2), flush=True)
    ctx.set_swat_trace_budget(0)
    try:
        a, b = gen.dna_pair(n1, n2, n1 + 7 * n2)
        for sc in [(2, -3, -5, -2), (10, -3, -3, -1)]:
            for tf in (0, 1):
                _check(ctx, a, b, sc, tf)
    finally:
        ctx.set_swat_trace_budget(32 << 30)


def test_checkpointed_related_sequences(ctx):
    """A long shared segment: the walk crosses many segments."""
    from dpx import gen
    base = gen.dna_uniform(6000, 5)
    a = base[:5000]
    b = gen.dna_uniform(300, 6) + base[800:4800] + gen.dna_uniform(200, 7)
    ctx.set_swat_trace_budget(0)
    try:
        for tf in (0, 1):
            _check(ctx, a, b, (2, -3, -5, -2), tf)
    finally:
        ctx.set_swat_trace_budget(32 << 30)


def test_ring_reuse_tall_alignment(ctx):
    """270000 x 60: 4219 bands, more than the ring's slots."""
    from dpx import gen
    a, b = gen.dna_pair(270000, 60, 11)
    print("swat 270000x60", flush=True)
    for tf in (0, 1):
        _check(ctx, a, b, (2, -3, -5, -2), tf)
    ctx.set_swat_trace_budget(0)
    try:
        _check(ctx, a, b, (2, -3, -5, -2), 1)
    finally:
        ctx.set_swat_trace_budget(32 << 30)


def _moves_consistent(tr):
    """Each record follows from the one before by the move its rule names."""
    step = {1: (-1, -1), 2: (0, 0), 3: (0, 0), 5: (-1, 0), 6: (-1, 0), 8: (0, -1), 9: (0, -1)}
    for p, q in zip(tr, tr[1:]):
        di, dj = step[p[2]]
        if (q[0], q[1]) != (p[0] + di, p[1] + dj):
            return False
    return tr[-1][2] in (0, 4, 7)


def test_checkpointed_matches_full_at_40kbp(ctx):
    """40000 x 36000 with a shared 30 kbp segment, beyond the oracle's reach
    here: the checkpointed walk equals the one through every code."""
    from dpx import gen
    base = gen.dna_uniform(40000, 21)
    a = base
    b = gen.dna_uniform(3000, 22) + base[5000:35000] + gen.dna_uniform(3000, 23)
    print("swat 40000x36000", flush=True)
    full = ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
    ctx.set_swat_trace_budget(0)
    try:
        cp = ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
    finally:
        ctx.set_swat_trace_budget(32 << 30)
    assert full[:2] == cp[:2]
    assert full[1][0] >= 2 * 30000 * 0.9  # the shared segment aligns
    assert [x[:3] for x in full[2]] == [x[:3] for x in cp[2]]
    assert _moves_consistent([x[:3] for x in cp[2]])


def test_checkpointed_matches_full_at_100kbp(ctx):
    """100000 x 100000 (1e10 cells; every code would take 5 GB): with a
    1 GiB budget the traceback recomputes segments and must equal the walk
    through every code."""
    from dpx import gen
    base = gen.dna_uniform(100000, 31)
    a = base
    b = gen.dna_uniform(5000, 32) + base[10000:90000] + gen.dna_uniform(15000, 33)
    print("swat 100000x100000 full", flush=True)
    full = ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
    ctx.set_swat_trace_budget(1 << 30)
    try:
        print("swat 100000x100000 checkpointed", flush=True)
        cp = ctx.swat(a, b, 2, -3, -5, -2, trace_from=1)
    finally:
        ctx.set_swat_trace_budget(32 << 30)
    assert full[:2] == cp[:2] and full[1][0] >= 2 * 80000 * 0.9
    assert full[2] == cp[2]
    assert _moves_consistent([x[:3] for x in cp[2]])


def test_size_and_score_limits(ctx):
    """The best cell travels as one 64-bit key (36-bit cell index, 28-bit
    score): larger problems are refused before any allocation."""
    import dpx
    with pytest.raises(dpx.DpxError, match="2\\^36"):
        ctx.swat("A" * 270000, "C" * 270000, 2, -3, -5, -2)
    with pytest.raises(dpx.DpxError, match="scores too large"):
        ctx.swat("A" * 1000, "A" * 1000, 1 << 20, -3, -5, -2)
