"""Writes tests/golden/rnafold_large.json: the oracle's RNAfold answers and
traces for inputs too long to recompute inside a GPU test (the oracle takes
about 20 s at n = 2048).  Inputs come from the seeded generators, so the file
holds only the generator call, the score and the trace.
Run:  python tests/golden/make_rnafold_large.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lamp-dp-mt_amd"))
sys.path.insert(0, os.path.dirname(HERE))
from dpx import gen  # noqa: E402
import oracle_lib as o  # noqa: E402

CASES = [("rna_gc", 2048, 2), ("rna_uniform", 1500, 9)]


def main():
    out = []
    for fn, n, seed in CASES:
        seq = getattr(gen, fn)(n, seed)
        score, tr = o.rnafold(seq)
        out.append({"gen": fn, "n": n, "seed": seed, "score": score, "trace": [list(t) for t in tr]})
        print(fn, n, seed, score, len(tr), flush=True)
    with open(os.path.join(HERE, "rnafold_large.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))


if __name__ == "__main__":
    main()
