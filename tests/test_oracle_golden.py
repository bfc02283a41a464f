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
