"""The reference's generated per-grammar C ABI (my_init / my_solve /
my_backtrack / my_free, CodeGen.scala:291-358) over libdpx.so.

tests/compat/harness.c is a harness written against that ABI, like
bench2/wrapper.c: it includes only the grammar's reference-shaped header
(include/compat/*.h) and is linked, with gcc, against the grammar's compat
translation unit (lamp-dp-mt_amd/compat/*.c, the replacement of the
generated C file) and libdpx.so.  Its printed scores and traces are compared
with the reference's golden vectors and the oracle.  The CPU tests only
build and link; the gpu ones run."""
import json
import os
import subprocess

import pytest

import oracle_lib as o

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
PKG = os.path.join(ROOT, "lamp-dp-mt_amd")
PAR = os.path.join(PKG, "params", "rna_turner2004.par")
ZU_TL = [2] * 6 + [1] * 9


def _load(name):
    with open(os.path.join(HERE, "golden", name)) as f:
        return json.load(f)


def _build(tmp, grammar, defs):
    """gcc harness.c + compat/<grammar>.c -ldpx; defs: extra -D flags."""
    macro = {"zu": "ZU", "nu": "NU", "rf": "RF", "mc": "MC", "sw": "SW"}[grammar]
    exe = os.path.join(str(tmp), "h_%s_%s" % (grammar, "_".join(d.replace("=", "") for d in defs)))
    if not os.path.exists(exe):
        cmd = ["gcc", "-std=gnu11", "-O2", "-Wall", "-Werror", "-D" + macro, '-DDPX_PARAMS="%s"' % PAR] + \
              ["-D" + d for d in defs] + \
              ["-I", os.path.join(ROOT, "include", "compat"), "-I", os.path.join(ROOT, "include"),
               os.path.join(HERE, "compat", "harness.c"), os.path.join(PKG, "compat", grammar + ".c"),
               "-L" + PKG, "-ldpx", "-Wl,-rpath," + PKG, "-o", exe]
        subprocess.check_call(cmd)
    return exe


def _run(exe, lines):
    out = subprocess.run([exe], input=("\n".join(lines) + "\n").encode(), capture_output=True, timeout=300)
    assert out.returncode == 0, out.stderr.decode()[-2000:]
    return [json.loads(x) for x in out.stdout.decode().strip().split("\n")]


@pytest.mark.parametrize("grammar", ["zu", "nu", "rf", "mc", "sw"])
def test_compat_builds_and_exports(tmp_path, grammar):
    """Each harness compiles with -Wall -Werror and links; the compat unit
    exports exactly the reference's four functions (and comp1.h's g_* for SW)."""
    exe = _build(tmp_path, grammar, ["SIZE=17"])
    syms = subprocess.run(["nm", "-g", "--defined-only", exe], capture_output=True, check=True).stdout.decode()
    names = {line.split()[-1] for line in syms.splitlines() if line.strip()}
    want = {"my_init", "my_solve", "my_backtrack", "my_free"}
    if grammar == "sw":
        want |= {"g_init", "g_solve", "g_backtrack", "g_free"}
    assert want <= names


@pytest.mark.skipif(not os.path.exists("/root/reference/bench2/wrapper.c"), reason="reference tree not mounted")
@pytest.mark.parametrize("kind", ["ZU", "NU"])
def test_reference_wrapper_links_unchanged(tmp_path, kind):
    """The reference's own harness (bench2/wrapper.c), compiled in place with
    its own header, links against the compat unit in place of the generated
    zu.c / nu.c (built and linked only: running it needs the GPU)."""
    unit = os.path.join(PKG, "compat", "zu.c" if kind == "ZU" else "nu.c")
    subprocess.check_call(["gcc", "-std=gnu11", "-O2", "-D" + kind, "-DSIZE=100", "-DLOOPS=1", "-DBACKTRACK", "-DCHECK",
                           "-I", os.path.join(ROOT, "include"), "/root/reference/bench2/wrapper.c", unit,
                           "-L" + PKG, "-ldpx", "-o", str(tmp_path / "wrapper")])


def _zn(tr):
    return [[a, b, r, p0, (p1 if ZU_TL[r] > 1 else 0)] for a, b, r, p0, p1 in tr]


@pytest.mark.gpu
def test_compat_zuker_golden(tmp_path):
    zu = _load("zuker_ref.json")
    pick = [f for f in zu if len(f["seq"]) in (13, 32, 100, 300, 500, 2048)]
    for f in pick:
        n = len(f["seq"])
        print("compat zuker n=%d" % n, flush=True)
        (got,) = _run(_build(tmp_path, "zu", ["SIZE=%d" % n]), [f["seq"]])
        assert got["score"] == f["score"]
        assert _zn(got["trace"]) == f["trace"]


@pytest.mark.gpu
def test_compat_nussinov_golden(tmp_path):
    nu = _load("nussinov_ref.json")
    for f in [f for f in nu if len(f["seq"]) in (5, 55, 500, 1024)]:
        n = len(f["seq"])
        print("compat nussinov n=%d" % n, flush=True)
        (got,) = _run(_build(tmp_path, "nu", ["SIZE=%d" % n]), [f["seq"]])
        assert got["score"] == f["score"]
        assert [[a, b, r, p0, 0] for a, b, r, p0, _ in got["trace"]] == f["trace"]


@pytest.mark.gpu
def test_compat_rnafold(tmp_path):
    from dpx import gen
    big = _load("rnafold_large.json")
    for f in big:
        n = f["n"]
        print("compat rnafold n=%d (fixture)" % n, flush=True)
        (got,) = _run(_build(tmp_path, "rf", ["SIZE=%d" % n]), [getattr(gen, f["gen"])(n, f["seed"])])
        assert got["score"] == f["score"]
        assert [t[:5] for t in got["trace"]] == [list(t[:5]) for t in f["trace"]]
    for n in (40, 150):
        seqs = [gen.rna_gc(n, 60 + n + k) for k in range(3)]
        outs = _run(_build(tmp_path, "rf", ["SIZE=%d" % n]), seqs)
        for s, got in zip(seqs, outs):
            osc, otr = o.rnafold(s)
            assert got["score"] == osc and [tuple(t[:4]) for t in got["trace"]] == [t[:4] for t in otr], n


@pytest.mark.gpu
def test_compat_matchain(tmp_path):
    from dpx import gen
    # Tests.scala:16-25: (10,7500,50) and (1,122,3)
    (a,) = _run(_build(tmp_path, "mc", ["SIZE=3"]), ["10 100 100 5 5 50"])
    assert a["score"] == 7500
    rows, cols = [1, 2, 20, 2, 4, 2, 1, 7], [2, 20, 2, 4, 2, 1, 7, 3]
    (b,) = _run(_build(tmp_path, "mc", ["SIZE=8"]), [" ".join("%d %d" % rc for rc in zip(rows, cols))])
    assert b["score"] == 122
    rows, cols = gen.matrix_chain(200, 77)
    (c,) = _run(_build(tmp_path, "mc", ["SIZE=200"]), [" ".join("%d %d" % rc for rc in zip(rows, cols))])
    osc, otr = o.matchain(rows, cols)
    assert c["score"] == osc and [tuple(t[:4]) for t in c["trace"]] == [t[:4] for t in otr]


@pytest.mark.gpu
@pytest.mark.parametrize("n1,n2", [(1, 1), (64, 63), (300, 1024), (1024, 1024)])
def test_compat_swat(tmp_path, n1, n2):
    """Two-track form (comp1.h): M_H-1 x M_W-1 sequences, SWatAlgebra's
    scores, trace from (n1, n2), records without positions."""
    from dpx import gen
    a, b = gen.dna_pair(n1, n2, 5000 + n1 + n2)
    print("compat swat %dx%d" % (n1, n2), flush=True)
    (got,) = _run(_build(tmp_path, "sw", ["M_H=%d" % (n1 + 1), "M_W=%d" % (n2 + 1)]), ["%s %s" % (a, b)])
    axiom, _, otr = o.swat(a, b, 10, -3, -3, -1)
    assert got["score"] == axiom
    assert [tuple(t[:3]) for t in got["trace"]] == [t[:3] for t in otr]
    assert got["trace"][0][:2] == [n1, n2]
