"""Deterministic target sampling for the benchmark configurations.

SplitMix64 (Steele, Lea, Flood, "Fast splittable pseudorandom number
generators", OOPSLA 2014) drives every draw, so the samples do not depend on
numpy's generator versions.

* C1: 4,096 distinct indices into the 7,633,612 pruned length-2 targets in Go
  order (seed 0), returned in ascending index order.
* C3: 65,536 length-3 targets drawn by rejection (seed 1): three independent
  uniform picks from the 2,863 instructions, kept when ShouldPrune is false
  (pkg/search/pruner.go:6-34) and not already drawn.
* C4/C5: 256 STOKE targets (seed 2): length uniform in 4..8, instructions
  uniform over the 2,863, rejected when a pull (PLA/PLP) meets an empty stack,
  i.e. the sequence reads a stack slot it never wrote.
"""
from __future__ import annotations

import numpy as np

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed: int):
        self.state = seed & MASK64

    def next(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def below(self, n: int) -> int:
        # Lemire-free simple rejection to stay exactly uniform.
        limit = (1 << 64) - ((1 << 64) % n)
        while True:
            v = self.next()
            if v < limit:
                return v % n


def sample_indices(n_total: int, k: int, seed: int) -> np.ndarray:
    rng = SplitMix64(seed)
    seen = set()
    while len(seen) < k:
        seen.add(rng.below(n_total))
    return np.array(sorted(seen), dtype=np.int64)


def c1_targets(pruned_l2: np.ndarray) -> np.ndarray:
    idx = sample_indices(len(pruned_l2), 4096, 0)
    return pruned_l2[idx].copy(), idx


def c3_targets(table: np.ndarray, should_prune, k: int = 65536, seed: int = 1) -> np.ndarray:
    rng = SplitMix64(seed)
    out, seen = [], set()
    n = len(table)
    while len(out) < k:
        tri = (rng.below(n), rng.below(n), rng.below(n))
        if tri in seen:
            continue
        seq = [int(table[tri[0]]), int(table[tri[1]]), int(table[tri[2]])]
        if should_prune(seq):
            continue
        seen.add(tri)
        out.append(seq)
    return np.array(out, dtype=np.uint32)


def c4_targets(table: np.ndarray, k: int = 256, seed: int = 2):
    rng = SplitMix64(seed)
    pla, plp, pha, php = 22, 24, 21, 23
    out = []
    n = len(table)
    while len(out) < k:
        length = 4 + rng.below(5)
        seq = [int(table[rng.below(n)]) for _ in range(length)]
        depth, ok = 0, True
        for w in seq:
            op = w & 0xFFFF
            if op in (pha, php):
                depth += 1
            elif op in (pla, plp):
                if depth == 0:
                    ok = False
                    break
                depth -= 1
        if ok:
            out.append(seq)
    return out
