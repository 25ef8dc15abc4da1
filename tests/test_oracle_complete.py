"""Pins the oracle's complete exhaustive check by projection (mode 2) and its
parts (CPU only).

Mode 2 decides verifier.go:262-415's predicate over the whole domain by
sweeping, per compared output, only the domain registers that output depends
on.  It is as sound as the oracle's dependency table (sxo_dep_units), which
is written from the executor; these tests check:

* the table: for every opcode, output unit and input unit outside the
  output's dependency set, every value of that input (all 256 of a byte,
  both of a flag bit) on random states with random immediates leaves the
  output unchanged;
* the block executor behind the brute force is the scalar executor;
* mode 2 against the brute force over the whole domain (mode 1) on pairs of
  every shape up to two extra registers, passing and failing, under several
  dead-flag masks (pairs with three extra registers are brute-forced in
  test_oracle_fallback.py).
"""
import os

import numpy as np
import pytest

import oracle as o

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BYTES = (0, 1, 2, 4, 5, 6, 7)  # state bytes other than P


def test_dependency_table_is_sound():
    rng = np.random.default_rng(21)
    nb = 384
    inputs = [(u, 256) for u in BYTES] + [(8 + b, 2) for b in range(8)]
    for op in range(58):
        dep = o.dep_units(op)
        imm = rng.integers(0, 256, nb) if 26 <= op <= 36 else np.zeros(nb, np.int64)
        base = rng.integers(0, 256, (nb, 8)).astype(np.uint8)
        for u, nv in inputs:
            st = np.repeat(base, nv, axis=0)
            vals = np.tile(np.arange(nv), nb)
            if u < 8:
                st[:, u] = vals
            else:
                b = u - 8
                st[:, 3] = (st[:, 3] & np.uint8(0xFF ^ (1 << b))) | (vals << b).astype(np.uint8)
            words = (op | (np.repeat(imm, nv) << 16)).astype(np.uint32)
            out = o.exec_many(st, words).reshape(nb, nv, 8)
            for q in list(BYTES) + [8 + b for b in range(8)]:
                if dep[q] & (1 << u):
                    continue
                v = out[..., q] if q < 8 else (out[..., 3] >> (q - 8)) & 1
                assert (v == v[:, :1]).all(), (o.OPS[op], "output", q, "depends on", u)


def test_block_executor_is_the_executor():
    rng = np.random.default_rng(22)
    st = rng.integers(0, 256, (512, 8)).astype(np.uint8)
    for w in o.instruction_table(0)[::7]:
        assert np.array_equal(o.exec_block(st, int(w)), o.exec_many(st, np.full(512, w, np.uint32)))


def _extras(t, c):
    reads = 0
    for w in t + c:
        reads |= o.op_reads(w & 0xFFFF)
    return sum(1 for b in (4, 8, 16, 32, 64, 128) if reads & b)


def _pairs():
    """Known rewrites of the length-2 fixtures (passing), the same with the
    candidate perturbed (mostly failing late), and hand-picked shapes."""
    W = o.word
    rng = np.random.default_rng(23)
    out = []
    for name, dead in (("c1_go_dead00", 0), ("c1_go_deadFF", 0xFF), ("c1_go_dead01", 0x01)):
        rules = np.load(os.path.join(GOLD, name + ".npy"), allow_pickle=False)
        for r in rules[rng.choice(len(rules), 24, replace=False)]:
            t = r["src"][: r["src_len"]].tolist()
            c = r["rep"][: r["rep_len"]].tolist()
            d = dead if r["pass_"] == 2 else 0
            out.append((t, c, d))
            w = c[0]
            c2 = [(w & 0xFFFF) | ((((w >> 16) + 1) & 0xFF) << 16)] if 26 <= (w & 0xFFFF) <= 36 else \
                [int(o.instruction_table(0)[rng.integers(0, 2863)])]
            out.append((t, c2, d))
    out += [([W(o.CPX_M), W(o.CMP_IMM, 9)], [W(o.CMP_IMM, 9)], 0), ([W(o.PHA), W(o.PLA)], [W(o.ORA_IMM, 0)], 0),
            ([W(o.TSX), W(o.INX)], [W(o.TSX)], 0xFF), ([W(o.INX), W(o.DEX)], [W(o.CPX_IMM, 0)], 0x01),
            ([W(o.DEX), W(o.CPY_IMM, 0x1D)], [W(o.DEX)], 0xFF), ([W(o.STX_M), W(o.LDA_M)], [W(o.STX_M), W(o.TXA)], 0),
            ([W(o.PHP), W(o.PLA)], [W(o.PHP), W(o.PLA)], 0), ([W(o.PHA), W(o.PLP)], [W(o.PHA), W(o.PLP)], 0x30)]
    return [p for p in out if _extras(p[0], p[1]) <= 2]


def test_projection_equals_brute_force():
    pairs = _pairs()
    assert len(pairs) >= 140
    seen = {True: 0, False: 0}
    for t, c, dead in pairs:
        brute, cex = o.exhaustive_brute(t, c, dead, nthreads=8)
        proj, pcex = o.exhaustive_cex(t, c, dead, o.EXH_COMPLETE)
        assert proj == brute, ([o.disasm(w) for w in t], [o.disasm(w) for w in c], dead)
        assert o.exhaustive_check(t, c, dead, o.EXH_GO) == brute  # <= 2 extra registers: Go sweeps fully
        seen[brute] += 1
        if not proj:  # the projection's counterexample is a real one
            a, b = o.exec_seq(list(pcex), t), o.exec_seq(list(pcex), c)
            assert any(x != y for k, (x, y) in enumerate(zip(a, b)) if k != 3) or (a[3] ^ b[3]) & 0xCF & ~dead
    assert seen[True] >= 60 and seen[False] >= 40
