"""dpx_zuker_noalloc (the allocation-free variant of dpx_zuker, on a static
per-device scratch) through the C ABI: the same MFE and trace as dpx_zuker on
two calls in a row, the second with a smaller input, the maximum size, and
the error beyond it."""
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    import dpx
    c = dpx.Context()
    yield c
    c.close()


def test_noalloc_equals_zuker_twice_then_smaller(ctx):
    from dpx import gen
    for n, seed in ((900, 31), (120, 32), (1024, 33), (0, 34), (1, 35), (33, 36)):
        s = gen.rna_gc(n, seed)
        print("zuker_noalloc n=%d" % n, flush=True)
        want = ctx.zuker(s)
        got = ctx.zuker_noalloc(s)
        assert got == want, n
        # parse only
        assert ctx.zuker_noalloc(s, trace=False)[0] == want[0]


def test_noalloc_golden(ctx):
    import json
    import os
    with open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "zuker_ref.json")) as f:
        zu = json.load(f)
    tl = [2] * 6 + [1] * 9
    for fx in zu:
        if len(fx["seq"]) > 1024:
            continue
        score, tr = ctx.zuker_noalloc(fx["seq"])
        assert score == fx["score"]
        assert [[a, b, r, p0, p1 if tl[r] > 1 else 0] for a, b, r, p0, p1 in tr] == fx["trace"]


def test_noalloc_rejects_beyond_maximum(ctx):
    import dpx
    from dpx import gen
    with pytest.raises(dpx.DpxError) as e:
        ctx.zuker_noalloc(gen.rna_gc(1025, 1))
    assert e.value.code == dpx.EINVAL
    assert "1024" in ctx._L.dpx_last_error(ctx.h).decode()
    # the context still works
    s = gen.rna_gc(64, 2)
    assert ctx.zuker_noalloc(s) == ctx.zuker(s)
