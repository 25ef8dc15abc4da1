"""Pins of the oracle's STOKE restatement (oracle/sx_oracle.c).

* Philox4x32-10 against the known-answer vectors of the Random123
  distribution (kat_vectors: zero, all-ones and pi-digit counters/keys).
* The proposal distribution's Metropolis-Hastings ratio: the closed form used
  by the chain equals q(x|x') / q(x'|x) computed by enumerating every
  proposal (move type, position, instruction, immediate) from both sides.
* The C chain equals a step-by-step Python restatement over hundreds of steps
  (uniform and non-uniform instruction weights).
* Cost identities of pkg/stoke/stoke_test.go:10-27,60-74 in this build's
  cost (10 x Hamming + bytes).
"""
import math
from collections import defaultdict

import numpy as np
import pytest

import oracle as o

W = o.word
TAB = o.instruction_table(0).tolist()
IDX = {w: i for i, w in enumerate(TAB)}
M32 = 1 << 32
WR, WS, WD, WI, WM = 102 / 256, 52 / 256, 51 / 256, 25 / 256, 26 / 256


def test_philox_known_answers():
    assert o.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]
    assert o.philox4x32_10([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2) == [0x408F276D, 0x41C83B0E, 0xA20BC7C6,
                                                                   0x6D5451FD]
    assert o.philox4x32_10([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0]) == \
        [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]


def uniform_weights():
    base, extra = divmod(M32, 2863)
    return np.array([base + (1 if i < extra else 0) for i in range(2863)], np.uint64)


def random_weights(seed):
    rng = np.random.default_rng(seed)
    p = 0.9 * rng.dirichlet(np.full(2863, 0.3)) + 0.1 / 2863   # a uniform floor, as the product mixes in
    w = np.floor(p * M32).astype(np.uint64)
    w[np.argmax(w)] += M32 - int(w.sum())
    return w


def has_imm(w):
    return 26 <= (w & 0xFFFF) <= 36


def proposals(x, q, max_len=10):
    """q(.|x) by enumerating every proposal; q is the per-instruction
    probability array.  Returns {tuple(seq): probability}."""
    out = defaultdict(float)
    L = len(x)
    # replace
    for p in range(L):
        for i in range(2863):
            y = list(x); y[p] = TAB[i]
            out[tuple(y)] += WR / L * q[i]
    # swap
    if L >= 2:
        for p in range(L - 1):
            y = list(x); y[p], y[p + 1] = y[p + 1], y[p]
            out[tuple(y)] += WS / (L - 1)
    else:
        out[tuple(x)] += WS
    # delete
    if L > 1:
        for p in range(L):
            out[tuple(x[:p] + x[p + 1:])] += WD / L
    else:
        out[tuple(x)] += WD
    # insert
    if L < max_len:
        for p in range(L + 1):
            for i in range(2863):
                out[tuple(x[:p] + [TAB[i]] + x[p:])] += WI / (L + 1) * q[i]
    else:
        out[tuple(x)] += WI
    # immediate
    pos = [p for p in range(L) if has_imm(x[p])]
    if pos:
        for p in pos:
            for v in range(256):
                y = list(x); y[p] = (x[p] & 0xFFFF) | (v << 16)
                out[tuple(y)] += WM / len(pos) / 256
    else:
        out[tuple(x)] += WM
    return out


def ratio_formula(x, y, q):
    """The closed form the chain uses for q(x|y) / q(y|x)."""
    qi = lambda w: q[IDX[w]]
    if len(y) == len(x) - 1:
        p = next(i for i in range(len(x)) if i == len(y) or x[i] != y[i])
        return WI * qi(x[p]) / WD
    if len(y) == len(x) + 1:
        p = next(i for i in range(len(y)) if i == len(x) or x[i] != y[i])
        return WD / (WI * qi(y[p]))
    diff = [i for i in range(len(x)) if x[i] != y[i]]
    if len(diff) == 2:
        return 1.0
    (p,) = diff
    L = len(x)
    nimm = sum(has_imm(w) for w in x)
    same = WM / (nimm * 256) if (x[p] & 0xFFFF) == (y[p] & 0xFFFF) and has_imm(x[p]) else 0.0
    return (WR * qi(x[p]) / L + same) / (WR * qi(y[p]) / L + same)


@pytest.mark.parametrize("weights", ["uniform", "learned"])
def test_proposal_ratio_is_exact(weights):
    q = (uniform_weights() if weights == "uniform" else random_weights(1)).astype(np.float64) / M32
    rng = np.random.default_rng(2)
    seqs = [[W(o.LDA_IMM, 5), W(o.TAX), W(o.LDA_IMM, 5)],
            [W(o.INX)],
            [TAB[i] for i in rng.integers(0, 2863, 10)],
            [W(o.CMP_IMM, 1), W(o.CMP_IMM, 2), W(o.PHA), W(o.PHA)]]
    for x in seqs:
        fwd = proposals(x, q)
        ys = [y for y in fwd if list(y) != x]
        for k in rng.choice(len(ys), size=min(25, len(ys)), replace=False):
            y = list(ys[k])
            back = proposals(y, q).get(tuple(x), 0.0)
            assert back > 0
            assert math.isclose(back / fwd[tuple(y)], ratio_formula(x, y, q), rel_tol=1e-9), (x, y)


def py_steps(t, nsteps, chain_id, seed, decay, w):
    """Step-by-step restatement of the chain in Python."""
    q = w.astype(np.float64) / M32
    cum = np.cumsum(w.astype(np.uint64))
    tb = sum(o.byte_size(x & 0xFFFF) for x in t)
    x, temp, cost10, acc = list(t), 1.0, o.stoke_cost10(t, t), 0
    max_len = max(len(t) + 2, 10)
    for s in range(nsteps):
        r = o.philox4x32_10([s & 0xFFFFFFFF, s >> 32, chain_id, 0], [seed & 0xFFFFFFFF, seed >> 32])
        a = o.philox4x32_10([s & 0xFFFFFFFF, s >> 32, chain_id, 1], [seed & 0xFFFFFFFF, seed >> 32])
        m = r[0] >> 24
        mh = lambda u, n: (u * n) >> 32
        pick = lambda u: TAB[int(np.searchsorted(cum, u, side="right"))]
        L = len(x)
        y = None
        if m < 102:
            p = mh(r[1], L); v = pick(r[2])
            if v != x[p]:
                y = list(x); y[p] = v
        elif m < 154:
            if L >= 2:
                p = mh(r[1], L - 1)
                if x[p] != x[p + 1]:
                    y = list(x); y[p], y[p + 1] = y[p + 1], y[p]
        elif m < 205:
            if L > 1:
                p = mh(r[1], L); y = x[:p] + x[p + 1:]
        elif m < 230:
            if L < max_len:
                p = mh(r[1], L + 1); v = pick(r[2]); y = x[:p] + [v] + x[p:]
        else:
            pos = [i for i in range(L) if has_imm(x[i])]
            if pos:
                p = pos[mh(r[1], len(pos))]
                v = (x[p] & 0xFFFF) | ((r[3] >> 24) << 16)
                if v != x[p]:
                    y = list(x); y[p] = v
        if y is not None:
            c10 = o.stoke_cost10(t, y)
            rt = ratio_formula(x, y, q)
            la = -((c10 - cost10) / 10.0) / temp + (math.log(rt) if rt > 0 else -math.inf)
            u = ((a[0] >> 5) * 67108864.0 + (a[1] >> 6) + 0.5) / 9007199254740992.0
            if la >= 0 or u < math.exp(la):
                x, cost10, acc = y, c10, acc + 1
        temp *= decay
    return x, cost10, acc


@pytest.mark.parametrize("weights", ["uniform", "learned"])
def test_chain_matches_python_restatement(weights):
    w = uniform_weights() if weights == "uniform" else random_weights(3)
    t = [W(o.LDA_IMM, 1), W(o.TAX), W(o.INX), W(o.CLC), W(o.ADC_IMM, 2)]
    for chain_id in (0, 7):
        ch = o.stoke_chain(t)
        o.stoke_steps(t, ch, 300, chain_id=chain_id, seed=42, decay=0.995,
                      weights=None if weights == "uniform" else w)
        x, cost10, acc = py_steps(t, 300, chain_id, 42, 0.995, w)
        assert list(ch.seq[: ch.len]) == x and ch.cost10 == cost10 and ch.accepted == acc


def test_cost_identities():
    # stoke_test.go:10-27,60-74 restated for the Hamming cost
    s = [W(o.INX)]
    assert o.stoke_cost10(s, s) == 10 * 0 + 1
    assert o.stoke_cost10([W(o.INX)], [W(o.INY)]) > 10
    assert o.stoke_cost10([W(o.LDA_IMM, 0x42)], [W(o.LDA_IMM, 0x42)], dead=0xFF) == 2
    # a flag-only difference costs nothing with all flags dead
    assert o.stoke_cost10([W(o.SEC), W(o.LDA_IMM, 0)], [W(o.LDA_IMM, 0)], dead=0xFF) == 2


def test_chain_never_empties_and_counts_steps():
    t = [W(o.TXA), W(o.TAX)]
    ch = o.stoke_chain(t)
    o.stoke_steps(t, ch, 2000, seed=9)
    assert ch.step == 2000 and 1 <= ch.len <= 10 and ch.accepted <= 2000
