"""Parameter-set switching on the device: the reference's other .par files
(src/librna/vienna/rna_turner1999.par, rna_andronescu2007.par,
dna_mathews2004.par) loaded with dpx_set_rna_params give, through the HIP
path, the reference's own Zuker answers and traces for those files (golden
fixtures from bench2/zu.c) and the oracle's RNAfold answers; switching back
restores the Turner 2004 results."""
import json
import os

import pytest

import oracle_lib as o

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
ZU_TL = [2] * 6 + [1] * 9
PSETS = ["turner1999", "andronescu2007", "dna_mathews2004"]


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


@pytest.mark.parametrize("pset", PSETS)
def test_parameter_file_zuker_and_rnafold(ctx, pset):
    import dpx
    from dpx import gen
    print("parameter file %s" % pset, flush=True)
    ctx.set_rna_params(dpx.PAR_FILES[pset])
    fx = _load("zuker_ref_%s.json" % pset)
    for f in fx:
        score, tr = ctx.zuker(f["seq"])
        assert score == f["score"] and _zn(tr) == f["trace"], (pset, len(f["seq"]))
    res, trs = ctx.zuker_batch([f["seq"] for f in fx])
    assert res == [f["score"] for f in fx]
    assert [_zn(t) for t in trs] == [f["trace"] for f in fx]
    for n, seed in [(37, 1), (150, 2), (333, 3)]:
        s = gen.rna_gc(n, 600 + seed)
        assert ctx.rnafold(s) == o.rnafold(s, par=pset), (pset, n)
    ctx.set_rna_params(dpx.PAR_TURNER2004)
    z = _load("zuker_ref.json")[3]
    score, tr = ctx.zuker(z["seq"])
    assert score == z["score"] and _zn(tr) == z["trace"]
