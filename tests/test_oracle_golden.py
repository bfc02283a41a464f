"""The oracle against the reference's own generated CPU code (bench2 zu.c /
nu.c), whose outputs are the committed golden fixtures."""
import json
import os

import pytest

import oracle_lib as o

HERE = os.path.dirname(os.path.abspath(__file__))
ZU_TL = [2] * 6 + [1] * 9


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def _norm_zu(tr):
    return [[a, b, r, p0, (p1 if ZU_TL[r] > 1 else 0)] for a, b, r, p0, p1 in tr]


@pytest.mark.parametrize("fx", _load("zuker_ref.json"), ids=lambda f: "n%d" % len(f["seq"]))
def test_zuker_matches_reference(fx):
    score, tr = o.zuker(fx["seq"])
    assert score == fx["score"]
    assert _norm_zu(tr) == fx["trace"]


@pytest.mark.parametrize("fx", _load("nussinov_ref.json"), ids=lambda f: "n%d" % len(f["seq"]))
def test_nussinov_matches_reference(fx):
    score, tr = o.nussinov(fx["seq"])
    assert score == fx["score"]
    assert [[a, b, r, p0, 0] for a, b, r, p0, _ in tr] == fx["trace"]


PSET_FIXTURES = [(p, f) for p in ("turner1999", "andronescu2007", "dna_mathews2004")
                 for f in _load("zuker_ref_%s.json" % p)]


@pytest.mark.parametrize("pset,fx", PSET_FIXTURES, ids=lambda x: x if isinstance(x, str) else "n%d" % len(x["seq"]))
def test_zuker_other_parameter_files_match_reference(pset, fx):
    """The same reference code reading the other .par files of its tree."""
    score, tr = o.zuker(fx["seq"], par=pset)
    assert score == fx["score"]
    assert _norm_zu(tr) == fx["trace"]


def test_parameter_files_differ():
    # the fixtures are not all the same answers under another name
    s = "GGGGAAAACCCCAUGCAUGCGGCAAAAGCC"
    assert len({o.zuker(s, trace=False, par=p)[0] for p in o.PAR_SETS}) >= 3
