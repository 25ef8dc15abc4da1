"""Pins the oracle's verifier, pruner, enumerator and search driver
(oracle/sx_oracle.c restating pkg/search/*.go).

* Known answers from pkg/search/search_test.go and verifier_test.go.
* The published counts: 7,633,612 pruned length-2 targets (README.md:132),
  2,863 instructions, 2,860 length-1 candidates surviving the pruner.
* The search driver (first hit, byte filter, flag-relaxed second pass) against
  a line-by-line Python loop over the oracle's QuickCheck/ExhaustiveCheck
  primitives, on sampled targets.
* ExhaustiveCheck against a brute-force sweep of the independent textbook
  model (tests/test_oracle_exec.py) on k <= 2 pairs.
"""
import os

import numpy as np
import pytest

import oracle as o
from test_oracle_exec import model

W = o.word


def seq(*items):
    out = []
    for it in items:
        if isinstance(it, tuple):
            out.append(W(it[0], it[1]))
        else:
            out.append(W(it))
    return out


def test_quickcheck_known_answers():  # search_test.go:24-39, :96-105
    s = seq((o.LDA_IMM, 0x42))
    assert o.quick_check(s, s)
    assert not o.quick_check(s, seq((o.LDA_IMM, 0x43)))
    assert not o.quick_check(seq((o.LDA_IMM, 0)), seq((o.EOR_IMM, 0)))


def test_known_rewrites():  # search_test.go:41-94
    assert o.exhaustive_check(seq(o.INX), seq(o.INX))
    for t, c in ((seq(o.TXA, o.TAX), seq(o.TXA)), (seq(o.CLC, o.CLC), seq(o.CLC)),
                 (seq((o.LDA_IMM, 0)), seq((o.AND_IMM, 0)))):
        assert o.quick_check(t, c) and o.exhaustive_check(t, c)
    assert o.exhaustive_check(seq(o.TXA, o.INX), seq(o.TXA, o.INX))


def test_fingerprint():  # search_test.go:107-122, verifier_test.go:9-23
    a = o.fingerprint(seq((o.LDA_IMM, 0x42)))
    assert a == o.fingerprint(seq((o.LDA_IMM, 0x42))) and len(a) == 64
    assert a != o.fingerprint(seq((o.LDA_IMM, 0x43)))
    fp = o.fingerprint(seq((o.ADC_IMM, 0x42)))
    assert fp == o.fingerprint(seq((o.ADC_IMM, 0x42)))
    assert all(fp[8 * i + 3] & 0x30 == 0 for i in range(8))


def test_vectors_and_flagdiff():  # verifier_test.go:25-61
    assert o.qc_vectors().shape == (8, 8) and o.mid_vectors().shape == (32, 8)
    assert np.array_equal(o.mid_vectors()[:8], o.qc_vectors())
    assert o.flag_diff(seq(o.INX), seq(o.INX)) == 0
    assert o.exhaustive_check(seq(o.ASL_A), seq(o.ASL_A), dead=0xFF)
    # a flag-only difference: SEC : LDA #0 vs LDA #0 differ in C only
    assert o.flag_diff(seq(o.SEC, (o.LDA_IMM, 0)), seq((o.LDA_IMM, 0))) == 0x01


def test_midcheck():  # search_test.go:177-190
    assert o.mid_check(seq(o.INX), seq(o.INX))
    assert not o.mid_check(seq(o.INX), seq(o.INY))


def test_pruner_cases():  # search_test.go:124-175
    assert o.should_prune(seq(o.NOP, o.INX))
    assert o.should_prune(seq(o.PHA, o.PHA, o.PHA))
    assert o.should_prune(seq(o.PLA))
    assert not o.should_prune(seq(o.PHA, o.PLA))
    assert not o.should_prune(seq(o.TXS, o.STA_M))
    assert o.should_prune(seq(o.STA_M, o.TXS))
    assert o.should_prune(seq((o.LDA_IMM, 1), (o.LDA_IMM, 2)))  # dead write of A


def test_published_counts():
    assert len(o.enumerate_pruned(1)) == 2860
    l2 = o.enumerate_pruned(2)
    assert len(l2) == 7_633_612                                   # README.md:132
    assert len(o.instruction_table()) == 2863


def _py_process_task(t, dead, cands):
    """worker.go:119-245 restated as a plain loop over the oracle primitives."""
    tb = sum(o.byte_size(w & 0xFFFF) for w in t)
    tc = sum(o.cycles(w & 0xFFFF) for w in t)
    rules = []
    for c in cands:
        if o.byte_size(c & 0xFFFF) >= tb or o.should_prune([c]):
            continue
        if o.quick_check(t, [c]) and o.exhaustive_check(t, [c]):
            rules.append((1, c, 0, tb - o.byte_size(c & 0xFFFF), tc - o.cycles(c & 0xFFFF)))
            break
    if dead:
        for c in cands:
            if o.byte_size(c & 0xFFFF) >= tb or o.should_prune([c]):
                continue
            if o.quick_check(t, [c]) or not o.quick_check(t, [c], dead):
                continue
            if not o.exhaustive_check(t, [c], dead):
                continue
            fd = o.flag_diff(t, [c])
            if fd == 0:
                continue
            rules.append((2, c, fd, tb - o.byte_size(c & 0xFFFF), tc - o.cycles(c & 0xFFFF)))
            break
    return rules


@pytest.mark.parametrize("dead", [0x00, 0xFF, 0x01])
def test_driver_matches_go_loop(dead):
    from sampling import SplitMix64
    l2 = o.enumerate_pruned(2)
    rng = SplitMix64(99)
    idx = sorted({rng.below(len(l2)) for _ in range(40)})
    # add targets known to have rules of both passes
    extra = [seq(o.CPX_M, (o.CMP_IMM, 9)), seq((o.LDA_IMM, 1), (o.EOR_IMM, 0xDC)),
             seq(o.DEX, (o.CPY_IMM, 0x1D)), seq(o.ROL_A, (o.AND_IMM, 0xFE)),
             seq(o.TXA, o.TAX), seq(o.CLC, o.CLC)]
    targets = np.array([l2[i].tolist() for i in idx] + extra, dtype=np.uint32)
    rules, st = o.search(targets, dead=dead, nthreads=4)
    cands = o.instruction_table(0).tolist()
    got = {}
    for r in rules:
        got.setdefault(int(r["target_index"]), []).append(
            (int(r["pass_"]), int(r["rep"][0]), int(r["dead_flags"]), int(r["bytes_saved"]),
             int(r["cycles_saved"])))
    for ti, t in enumerate(targets.tolist()):
        assert got.get(ti, []) == _py_process_task(t, dead, cands), [o.disasm(w) for w in t]


def test_readme_example_rules():
    """README.md:14-16,145-156 examples under both byte models."""
    def first(t, **kw):
        rules, _ = o.search(np.array([t], dtype=np.uint32), **kw)
        return [(o.disasm(int(r["rep"][0])), int(r["bytes_saved"]), int(r["cycles_saved"]))
                for r in rules]
    assert first(seq(o.TXA, o.TAX)) == [("TXA", 1, 2)]
    assert first(seq(o.ROL_A, (o.AND_IMM, 0xFE)))[0][0] == "ASL A"
    assert first(seq(o.ROR_A, (o.AND_IMM, 0x7F)))[0][0] == "LSR A"
    assert first(seq(o.BIT_M, (o.ADC_IMM, 0)))[0][0] == "ADC #$00"
    assert first(seq(o.LDA_M, o.STA_M))[0][0] == "LDA M"
    # LDA M : EOR M -> LDA #$00 exists only with the CUDA tool's 2-byte M ops
    assert first(seq(o.LDA_M, o.EOR_M)) == []
    cuda = first(seq(o.LDA_M, o.EOR_M), bytes_model=1, hits=1)
    assert ("LDA #$00", 2, 4) in cuda


def _brute_chunk(args):
    """States [lo, hi) of the sweep grid through the textbook model: True when
    target and candidate agree on all of them."""
    t, c, dead, extra, lo, hi = args
    idx = np.arange(lo, hi, dtype=np.int64)
    st = np.zeros((hi - lo, 8), np.uint8)
    st[:, 4] = 0xFD
    st[:, 0] = idx & 0xFF
    st[:, 3] = (idx >> 8) & 1
    for j, b in enumerate(extra):
        st[:, b] = (idx >> (9 + 8 * j)) & 0xFF
    ts, cs = st, st
    for w in t:
        ts = model(w & 0xFFFF, ts, np.full(len(ts), w >> 16))
    for w in c:
        cs = model(w & 0xFFFF, cs, np.full(len(cs), w >> 16))
    pm = 0xCF & ~dead
    diff = (ts != cs)
    diff[:, 3] = (ts[:, 3] & pm) != (cs[:, 3] & pm)
    return not diff.any()


def _brute_exh(t, c, dead):
    """Complete sweep of the domain with the textbook model (k <= 2): A (256)
    x carry (2) x each extra byte read (256 each), S = 0xFD unless read;
    chunks of 2^21 states spread over the host's cores."""
    import multiprocessing as mp
    reads = 0
    for w in t + c:
        reads |= o.op_reads(w & 0xFFFF)
    extra = [b for r, b in ((o.REG_X, 1), (o.REG_Y, 2), (o.REG_M, 5), (o.REG_S, 4),
                            (o.REG_S0, 6), (o.REG_S1, 7)) if reads & r]
    assert len(extra) <= 2
    n = 512 << (8 * len(extra))
    step = 1 << 21
    jobs = [(list(t), list(c), dead, extra, lo, min(n, lo + step)) for lo in range(0, n, step)]
    if len(jobs) == 1:
        return _brute_chunk(jobs[0])
    with mp.get_context("fork").Pool(min(8, len(jobs), os.cpu_count() or 1)) as pool:
        return all(pool.imap_unordered(_brute_chunk, jobs))


@pytest.mark.parametrize("t,c,dead", [
    (seq(o.CPX_M, (o.CMP_IMM, 9)), seq((o.CMP_IMM, 9)), 0),
    pytest.param(seq((o.CPX_IMM, 5), (o.CPY_IMM, 0xE5)), seq((o.CPY_IMM, 0xE5)), 0),
    pytest.param(seq(o.STX_M, o.LDA_M), seq(o.STX_M, o.TXA), 0),
    (seq(o.PHA, o.PLA), seq((o.ORA_IMM, 0)), 0),
    pytest.param(seq(o.DEX, (o.CPY_IMM, 0x1D)), seq(o.DEX), 0xFF),
    (seq(o.TSX, o.INX), seq(o.TSX), 0xFF),
    (seq(o.INX, o.DEX), seq((o.CPX_IMM, 0)), 0x01),
])
def test_exhaustive_vs_textbook_bruteforce(t, c, dead):
    assert o.exhaustive_check(t, c, dead) == _brute_exh(t, c, dead)
    assert o.exhaustive_check(t, c, dead, mode=1) == _brute_exh(t, c, dead)
