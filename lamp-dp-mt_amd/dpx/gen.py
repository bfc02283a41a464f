"""Seeded synthetic inputs for the DP workloads.

A self-contained splitmix64 stream, so that every sequence set is the same
on any machine and numpy version.  Generators:
  rna_uniform(n, seed)        i.i.d. uniform over ACGU
  rna_gc(n, seed, gc=0.5)     i.i.d. with P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2
  dna_uniform(n, seed)        i.i.d. uniform over ACGT
  log_uniform_lengths(k, lo, hi, seed)
  matrix_chain(n, seed, lo=1, hi=1000)  consistent (rows, cols) pairs
"""
import math

import numpy as np

MASK = (1 << 64) - 1
GAMMA = 0x9E3779B97F4A7C15


def _mix(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
    return z ^ (z >> 31)


class SplitMix64:
    def __init__(self, seed):
        # the seed is hashed, so nearby seeds start unrelated streams
        self.state = _mix((seed + 0x632BE59BD9B4E019) & MASK)

    def next(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        return _mix(self.state)

    def uniform(self):
        return (self.next() >> 11) * (1.0 / (1 << 53))

    def below(self, m):
        return self.next() % m


def _stream(seed, n):
    """The first n outputs of SplitMix64(seed).next(), vectorised (uint64
    arithmetic wraps modulo 2**64 exactly like the scalar class)."""
    s0 = np.uint64(_mix((seed + 0x632BE59BD9B4E019) & MASK))
    with np.errstate(over="ignore"):
        z = s0 + np.arange(1, n + 1, dtype=np.uint64) * np.uint64(GAMMA)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def _letters(alphabet, idx):
    return np.frombuffer(alphabet.encode(), dtype=np.uint8)[idx].tobytes().decode()


def rna_uniform(n, seed):
    return _letters("ACGU", (_stream(seed, n) % np.uint64(4)).astype(np.int64))


def rna_gc(n, seed, gc=0.5):
    """P(G)=P(C)=gc/2, P(A)=P(U)=(1-gc)/2 from one uniform draw per base."""
    u = (_stream(seed, n) >> np.uint64(11)).astype(np.float64) * (1.0 / (1 << 53))
    idx = np.where(u < gc / 2, 0, np.where(u < gc, 1, np.where(u < gc + (1 - gc) / 2, 2, 3)))
    return _letters("GCAU", idx)


def dna_uniform(n, seed):
    return _letters("ACGT", (_stream(seed, n) % np.uint64(4)).astype(np.int64))


def dna_pair(n1, n2, seed):
    """Two sequences drawn back to back from one seeded stream."""
    s = dna_uniform(n1 + n2, seed)
    return s[:n1], s[n1:]


def log_uniform_lengths(k, lo, hi, seed):
    r = SplitMix64(seed)
    a, b = math.log(lo), math.log(hi)
    return [min(hi, max(lo, int(round(math.exp(a + (b - a) * r.uniform()))))) for _ in range(k)]


def rna_batch(k, lo, hi, seed, gc=0.5):
    """k sequences with log-uniform lengths in [lo,hi] (the batched-fold set)."""
    lens = log_uniform_lengths(k, lo, hi, seed)
    return [rna_gc(n, seed * 1000003 + q + 1, gc) for q, n in enumerate(lens)]


def matrix_chain(n, seed, lo=1, hi=1000):
    r = SplitMix64(seed)
    dims = [lo + r.below(hi - lo + 1) for _ in range(n + 1)]
    return dims[:-1], dims[1:]
