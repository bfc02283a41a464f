"""Writes the fixtures of the batched and served Zuker workloads from the
reference's own generated CPU code (bench2/zu.c + zu_kern.h with src/librna,
built from the reference tree by oracle/Makefile into oracle/_ref/).  Run
here, where the reference is mounted:  python tests/golden/make_configs.py

zuker_batch4096_sample.json -- a stratified sample of the batched workload
  (4096 sequences, lengths log-uniform [50,2000], 50% GC, seed 3): in every
  length bucket (n in (32<<b, 64<<b]) the shortest, the longest and two
  between, plus the 8 longest sequences of the batch.
zuker_served_sample.json -- the Zuker requests of the first 2 s of the
  served stream (300 requests/s, seed 5, lengths log-uniform 30-3000): three
  per length bucket plus the 4 longest.
Each entry holds the request's index in its set, the sequence, and the
reference's MFE and trace (i, j, rule, pos0, pos1).
"""
import json
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "lamp-dp-mt_amd"))
from dpx import gen  # noqa: E402

ORACLE = os.path.join(ROOT, "oracle")


def bucket(n):
    b = 0
    while b < 7 and n > (64 << b):
        b += 1
    return b


def stratified(lengths, per_bucket, longest):
    """Indices: per length bucket `per_bucket` requests spread over the
    bucket's sorted lengths (its shortest and longest included), plus the
    `longest` longest overall."""
    by_b = {}
    for q, n in enumerate(lengths):
        by_b.setdefault(bucket(n), []).append(q)
    pick = set()
    for qs in by_b.values():
        qs = sorted(qs, key=lambda q: (lengths[q], q))
        m = len(qs)
        for k in range(per_bucket):
            pick.add(qs[round(k * (m - 1) / max(1, per_bucket - 1))])
    pick.update(sorted(range(len(lengths)), key=lambda q: (-lengths[q], q))[:longest])
    return sorted(pick)


def run_ref_parallel(seqs, workers=8):
    lens = sorted({len(s) for s in seqs})
    targets = ["_ref/zu_%d" % n for n in lens]
    subprocess.check_call(["make", "-s", "-j8", "-C", ORACLE] + targets)

    def one(n):
        group = [s for s in seqs if len(s) == n]
        res = subprocess.run([os.path.join(ORACLE, "_ref/zu_%d" % n)], input=("\n".join(group) + "\n").encode(),
                             capture_output=True, check=True)
        return {s: json.loads(line) for s, line in zip(group, res.stdout.decode().strip().split("\n"))}

    out = {}
    # longest first, so the pool's tail is short
    with ThreadPoolExecutor(workers) as ex:
        for d in ex.map(one, sorted(lens, reverse=True)):
            out.update(d)
    return out


def main():
    batch = gen.rna_batch(4096, 50, 2000, 3, gc=0.5)
    b_idx = stratified([len(s) for s in batch], 4, 8)
    stream = gen.served_stream(300.0, 2.0, 5)
    zq = [q for q, r in enumerate(stream) if r[1] == "zuker"]
    s_pick = stratified([len(stream[q][2]) for q in zq], 3, 4)
    s_idx = [zq[k] for k in s_pick]
    seqs = [batch[q] for q in b_idx] + [stream[q][2] for q in s_idx]
    print("reference runs: %d sequences, longest %d" % (len(seqs), max(len(s) for s in seqs)), flush=True)
    ref = run_ref_parallel(seqs)
    for name, idx, src in (("zuker_batch4096_sample.json", b_idx, lambda q: batch[q]),
                           ("zuker_served_sample.json", s_idx, lambda q: stream[q][2])):
        rows = [dict(index=q, seq=src(q), score=ref[src(q)]["score"], trace=ref[src(q)]["trace"]) for q in idx]
        with open(os.path.join(HERE, name), "w") as f:
            json.dump(rows, f, separators=(",", ":"))
        print("%s: %d entries, lengths %s" % (name, len(rows), [len(r["seq"]) for r in rows]))


if __name__ == "__main__":
    main()
