"""Writes the golden fixtures of tests/golden/ from the reference's own
generated CPU code (bench2/zu.c, bench2/nu.c), built from the reference tree
by oracle/Makefile into oracle/_ref/.  Run here, where the reference is
mounted:  python tests/golden/make_golden.py
Each fixture holds the input sequence and the reference's score and trace
(i, j, rule, pos0, pos1; pos entries the rule does not use are 0).
"""
import json
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lamp-dp-mt_amd"))
from dpx import gen  # noqa: E402

ORACLE = os.path.join(ROOT, "oracle")

# Zuker test sequences of the reference's smoke suite (Tests.scala:34-47)
ZUKER_TESTS = [
    "ccuuuuucaaagg", "guacgucaguacguacgugacugucagucaac", "aaaaaggaaacuccucuuu",
    "uucuccucaggaaga",
    "aaaaaagggaaaagaacaaaggagacucuucuccuuuuucaaaggaagaggagacucuuucaaaaaucccucuuuu",
    "gccaaccucgugca", "ggccaaccucgugcaa", "guugcucagcacgcguaaga",
    "gggcgcucaaccgagucagcagugcaauauagggccc",
    "augggcgcucaacucuccgugaauuugaaugagucagcagugcaauauagggcccucauc",
    "accacuccucauuugacuuauaggcucagaauuaguagaccacaguucacugugaaagga",
    "uugcccuaugucaaacauaugucgcaaagcacacgucguauucaccacgaucaaccaggg",
    "ccgaugccagcgucugcgccuucgccuaagggggagaagaagcucucccauaacggcaug",
    "ugcucaggca",
    # commented-out ones (Tests.scala:46-53), still valid inputs
    "agccccgguuaagaauaaaggagauuucuccgcccaaccccuguaaugcu",
    "ccggcgcccauaaaaucaaauuaacaucguuaugucagcaaguguaccacaagcuggaga",
]


# other parameter files of the reference (src/librna/vienna/*.par), read by
# the same reference code (bench2/zu.c built per set by oracle/Makefile, PSET=...)
PARAM_SETS = ["turner1999", "andronescu2007", "dna_mathews2004"]


def run_ref(kind, seqs, pset=None):
    by_len = {}
    for s in seqs:
        by_len.setdefault(len(s), []).append(s)
    out = {}
    for n, group in sorted(by_len.items()):
        target = "_ref/%s_%d" % (kind, n) if pset is None else "_ref/p_%s/%s_%d" % (pset, kind, n)
        exe = os.path.join(ORACLE, target)
        if not os.path.exists(exe):
            cmd = ["make", "-s", "-C", ORACLE, target] + ([] if pset is None else ["PSET=" + pset])
            subprocess.check_call(cmd)
        res = subprocess.run([exe], input=("\n".join(group) + "\n").encode(), capture_output=True, check=True)
        for s, line in zip(group, res.stdout.decode().strip().split("\n")):
            out[s] = json.loads(line)
    return [dict(seq=s.upper(), score=out[s]["score"], trace=out[s]["trace"]) for s in seqs]


def main():
    zs = [s.upper() for s in ZUKER_TESTS]
    for k, n in enumerate([7, 8, 9, 12, 16, 20, 25, 31, 40, 50, 64, 77, 100, 128, 150, 200, 257, 300, 400, 500]):
        zs.append(gen.rna_gc(n, 100 + k))
    for k, n in enumerate([60, 90, 120, 180, 240]):
        zs.append(gen.rna_gc(n, 200 + k, gc=0.7))
    zs.append(gen.rna_gc(2048, 2))  # the n=2048 MFE workload (50% GC, seed 2)
    z = run_ref("zu", zs)
    with open(os.path.join(HERE, "zuker_ref.json"), "w") as f:
        json.dump(z, f, separators=(",", ":"))
    ns = []
    for k, n in enumerate([1, 2, 3, 5, 8, 13, 21, 34, 55, 89, 100, 144, 233, 377, 500]):
        ns.append(gen.rna_uniform(n, 300 + k))
    ns.append(gen.rna_uniform(1024, 0))  # the n=1024 Nussinov workload (seed 0)
    nu = run_ref("nu", ns)
    with open(os.path.join(HERE, "nussinov_ref.json"), "w") as f:
        json.dump(nu, f, separators=(",", ":"))
    print("zuker fixtures: %d, nussinov fixtures: %d" % (len(z), len(nu)))
    param_sets()


def param_sets():
    ps = [s.upper() for s in ZUKER_TESTS[:8]]
    for k, n in enumerate([7, 12, 20, 33, 64, 100, 150, 257, 400]):
        ps.append(gen.rna_gc(n, 500 + k))
    ps.append(gen.rna_gc(120, 520, gc=0.7))
    ps.append(gen.rna_gc(800, 521))
    for pset in PARAM_SETS:
        z = run_ref("zu", ps, pset)
        with open(os.path.join(HERE, "zuker_ref_%s.json" % pset), "w") as f:
            json.dump(z, f, separators=(",", ":"))
        print("zuker fixtures with %s: %d" % (pset, len(z)))


if __name__ == "__main__":
    # "params": only the fixtures of the other parameter files
    param_sets() if sys.argv[1:] == ["params"] else main()
