"""Learned STOKE proposals: harvest rewrites, train the proposal network on
the device, and turn its predictions into the instruction tables that
sx_stoke_run samples from.

The reference draws every instruction uniformly (pkg/stoke/mutator.go:108-121,
randomInstruction).  Here a target's table is the network's distribution
given the target, mixed with a uniform share ``eps`` so that every
instruction stays proposable (the chain stays irreducible and every proposal
ratio finite); the Metropolis-Hastings acceptance in the kernel uses the
exact ratio of these tables, so the chain's stationary distribution is the
same as under uniform proposals -- only how fast it gets there changes.

Training examples are (context, label): the context is the rewritten
sequence as eight tokens (Go instruction index, 2863 = padding); the label is
an output of the network -- an instruction index, or 2863 + j, the pointer
to context position j.  An instruction of the sequence is labelled by its
pointer, so the network learns which position to act on rather than having to
reproduce an immediate.  ``labels="dropped"`` labels the instructions the
verified replacement drops (the deletions a chain must accept, whose
acceptance ratio is W_INS * q(v) / W_DEL), ``"rewrite"`` the replacement's
instructions (what insert and replace moves must propose), ``"both"`` either.
Sources are the enumerator's rules (length 2 and 3 targets) and
the rewrites STOKE chains verify on sequences drawn like the evaluation
targets but from other seeds.
"""
from __future__ import annotations

from collections import Counter
from typing import Iterable, List, Optional, Sequence, Tuple

import numpy as np

from . import core, isa

PAD = core.PAD_TOKEN
TABLE = np.array(isa.instruction_table(), dtype=np.uint32)
INDEX = {int(w): i for i, w in enumerate(TABLE)}
_PHA, _PHP, _PLA, _PLP = isa.OP["PHA"], isa.OP["PHP"], isa.OP["PLA"], isa.OP["PLP"]

Pair = Tuple[List[int], List[int]]


def tokens(seqs: Sequence[Sequence[int]]) -> np.ndarray:
    """Instruction words -> n x 8 tokens (Go index; 2863 pads the tail)."""
    out = np.full((len(seqs), core.MLP_CTX), PAD, dtype=np.uint16)
    for i, s in enumerate(seqs):
        if len(s) > core.MLP_CTX:
            raise ValueError("context holds at most 8 instructions")
        for k, w in enumerate(s):
            out[i, k] = INDEX[int(w)]
    return out


def stack_safe(seq: Sequence[int]) -> bool:
    """False when the sequence pulls from a stack slot it did not push."""
    depth = 0
    for w in seq:
        op = int(w) & 0xFFFF
        if op in (_PHA, _PHP):
            depth += 1
        elif op in (_PLA, _PLP):
            if depth == 0:
                return False
            depth -= 1
    return True


def random_targets(k: int, seed: int, min_len: int = 4, max_len: int = 8) -> List[List[int]]:
    """Sequences of min_len..max_len instructions drawn uniformly from the
    2,863, rejected when they read an undefined stack slot."""
    rng = np.random.default_rng(seed)
    out: List[List[int]] = []
    while len(out) < k:
        n = int(rng.integers(min_len, max_len + 1))
        seq = [int(w) for w in TABLE[rng.integers(0, len(TABLE), size=n)]]
        if stack_safe(seq):
            out.append(seq)
    return out


def rule_pairs(rules: np.ndarray) -> List[Pair]:
    """(source, replacement) word lists of enumerator rules (RULE_DTYPE)."""
    return [([int(w) for w in r["src"][:r["src_len"]]], [int(w) for w in r["rep"][:r["rep_len"]]])
            for r in rules]


def stoke_pairs(targets: Sequence[Sequence[int]], results: np.ndarray) -> List[Pair]:
    """(target, verified replacement) of every target STOKE verified."""
    return [(list(map(int, t)), [int(w) for w in r["rep"][:r["rep_len"]]])
            for t, r in zip(targets, results) if r["verified"]]


def examples(pairs: Iterable[Pair], labels: str = "both", repeat: int = 1) -> Tuple[np.ndarray, np.ndarray]:
    """Training examples of rewrite pairs: one per instruction the
    replacement drops (multiset difference source - replacement) and/or one
    per replacement instruction, as `labels` selects."""
    if labels not in ("both", "rewrite", "dropped"):
        raise ValueError(labels)
    ctx_rows: List[List[int]] = []
    lab: List[int] = []
    for src, rep in pairs:
        used = [False] * len(src)

        def label_of(w: int) -> int:
            for j, x in enumerate(src):
                if x == w and not used[j]:
                    used[j] = True
                    return core.MLP_CLASSES + j
            return INDEX[w]

        ys: List[int] = []
        if labels != "rewrite":
            dropped = Counter(src) - Counter(rep)
            ys += [label_of(w) for w in dropped.elements()]
        if labels != "dropped":
            ys += [label_of(w) for w in rep]
        for y in ys:
            for _ in range(repeat):
                ctx_rows.append(src)
                lab.append(y)
    if not lab:
        return np.zeros((0, core.MLP_CTX), np.uint16), np.zeros(0, np.uint16)
    return tokens(ctx_rows), np.array(lab, dtype=np.uint16)


def harvest_enumerator(dev: core.Device, l3_targets: Optional[np.ndarray] = None, dead: int = 0,
                       l3_dead: int = 0x01) -> List[Pair]:
    """Rules of the full length-2 search (dead flags `dead`) and of the given
    length-3 targets (dead flags `l3_dead`)."""
    pairs = rule_pairs(dev.enumerate(None, target_len=2, max_cand_len=1, dead=dead)[0])
    if l3_targets is not None and len(l3_targets):
        pairs += rule_pairs(dev.enumerate(np.asarray(l3_targets, np.uint32), target_len=3, max_cand_len=2,
                                          dead=l3_dead)[0])
    return pairs


def harvest_stoke(dev: core.Device, targets: Sequence[Sequence[int]], *, chains: int = 4, steps: int = 20000,
                  t0: float = 0.01, decay: float = 1.0, seed: int = 1, weights: Optional[np.ndarray] = None,
                  dead: int = 0) -> List[Pair]:
    """Verified rewrites STOKE chains find on the given targets (by default
    cold chains: at T = 0.01 a chain started on its target keeps every
    output and accepts the deletions that shorten it)."""
    res = dev.stoke(targets, chains=chains, steps=steps, t0=t0, decay=decay, seed=seed, dead=dead,
                    weights=weights, stop_when_verified=True)
    return stoke_pairs(targets, res)


def train_proposal(dev: core.Device, ctx: np.ndarray, labels: np.ndarray, *, steps: int = 20000,
                   batch: int = 1024, lr: float = 2e-3, seed: int = 0) -> Tuple[core.Mlp, np.ndarray]:
    """Train a fresh network for `steps` Adam steps on the examples."""
    if len(labels) == 0:
        raise ValueError("no training examples")
    m = dev.mlp(seed=seed)
    losses = m.train(ctx, labels, steps=steps, batch=batch, lr=lr, seed=seed + 1)
    return m, losses


def proposal_tables(m: core.Mlp, targets: Sequence[Sequence[int]], eps: float = 0.05) -> np.ndarray:
    """Per-target uint32 instruction tables (rows sum to 2^32) for sx_stoke_run."""
    return m.proposal(tokens(targets), eps)
