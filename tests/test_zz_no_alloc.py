"""sx_check_pairs_noalloc against sx_check_pairs, through the C ABI.

The no-alloc entry point keeps its device buffers in one fixed array per GPU
instead of allocating them; its result bits must equal sx_check_pairs' for
the same pairs (bit-exact), call after call, also when a later call is
smaller than an earlier one (the array keeps the earlier call's contents).
"""
import os

import numpy as np
import pytest

import oracle as o
from sx6502 import core

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
ALL = core.STAGE_QC | core.STAGE_MID | core.STAGE_EXH


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


def _pairs(seed: int, n: int):
    """n (target, candidate) pairs of up to 4 words: known rewrites of the
    length-2 fixture (which pass every stage), the same with the candidate's
    immediate or opcode perturbed (which mostly fail late), and random
    sequences of length 1-3 (which mostly fail QuickCheck)."""
    rng = np.random.default_rng(seed)
    rules = np.concatenate([np.load(os.path.join(GOLD, f), allow_pickle=False)
                            for f in ("c1_go_deadFF.npy", "c1_go_dead00.npy")])
    tab = o.instruction_table(0)
    tw = np.zeros((n, 4), np.uint32)
    cw = np.zeros((n, 4), np.uint32)
    tl = np.zeros(n, np.uint8)
    cl = np.zeros(n, np.uint8)
    for i in range(n):
        kind = rng.integers(0, 3)
        if kind < 2:
            r = rules[rng.integers(0, len(rules))]
            tw[i, :r["src_len"]] = r["src"][: r["src_len"]]
            cw[i, :r["rep_len"]] = r["rep"][: r["rep_len"]]
            tl[i], cl[i] = r["src_len"], r["rep_len"]
            if kind == 1:
                j = rng.integers(0, cl[i])
                w = int(cw[i, j])
                cw[i, j] = (w & 0xFFFF) | (((w >> 16) ^ int(rng.integers(1, 256))) << 16) if 26 <= (w & 0xFFFF) <= 36 \
                    else int(tab[rng.integers(0, len(tab))])
        else:
            tl[i], cl[i] = rng.integers(1, 4), rng.integers(1, 4)
            tw[i, :tl[i]] = tab[rng.integers(0, len(tab), tl[i])]
            cw[i, :cl[i]] = tab[rng.integers(0, len(tab), cl[i])]
        # garbage beyond each length: both entry points must ignore it
        tw[i, tl[i]:] = rng.integers(0, 1 << 20, 4 - tl[i])
        cw[i, cl[i]:] = rng.integers(0, 1 << 20, 4 - cl[i])
    return (tw, tl), (cw, cl)


def test_noalloc_equals_check_pairs_on_two_calls(dev):
    print("no-alloc: call 1 (6,000 pairs, dead 0xFF), call 2 (1,500 pairs, dead 0x00)", flush=True)
    t1, c1 = _pairs(1, 6000)
    a1 = dev.check_pairs_noalloc(t1, c1, ALL, dead=0xFF)
    b1 = dev.check_pairs(t1, c1, ALL, dead=0xFF)
    assert np.array_equal(a1, b1)
    assert (a1 & core.STAGE_EXH).sum() > 1000 and (a1 == 0).sum() > 100  # both outcomes are exercised
    t2, c2 = _pairs(2, 1500)
    a2 = dev.check_pairs_noalloc(t2, c2, ALL, dead=0x00)
    b2 = dev.check_pairs(t2, c2, ALL, dead=0x00)
    assert np.array_equal(a2, b2)
    # stage subsets and the full-domain sweep, again smaller
    a3 = dev.check_pairs_noalloc((t2[0][:700], t2[1][:700]), (c2[0][:700], c2[1][:700]), core.STAGE_QC, dead=0x01)
    b3 = dev.check_pairs((t2[0][:700], t2[1][:700]), (c2[0][:700], c2[1][:700]), core.STAGE_QC, dead=0x01)
    assert np.array_equal(a3, b3)
    # full-domain sweeps only where they stay at most 2^33 states
    keep = [i for i in range(len(t2[1])) if core.exh_plan(t2[0][i, :t2[1][i]].tolist(), c2[0][i, :c2[1][i]].tolist(),
                                                           0, core.EXH_FULL)[1] <= 2 ** 33][:300]
    small = (t2[0][keep], t2[1][keep]), (c2[0][keep], c2[1][keep])
    a4 = dev.check_pairs_noalloc(*small, core.STAGE_EXH, dead=0x00, exh_mode=core.EXH_FULL)
    b4 = dev.check_pairs(*small, core.STAGE_EXH, dead=0x00, exh_mode=core.EXH_FULL)
    assert np.array_equal(a4, b4)


def test_noalloc_known_answers(dev):
    W = o.word
    t = [[W(o.TXA), W(o.TAX)], [W(o.CLC), W(o.CLC)], [W(o.LDA_IMM, 0)], [W(o.LDA_IMM, 0)]]
    c = [[W(o.TXA)], [W(o.CLC)], [W(o.AND_IMM, 0)], [W(o.EOR_IMM, 0)]]
    got = dev.check_pairs_noalloc(t, c, ALL)
    assert got.tolist() == [ALL, ALL, ALL, 0]  # search_test.go:48-105


def test_noalloc_rejects_oversize_and_empty_is_ok(dev):
    n = core.NOALLOC_MAX_PAIRS + 1
    tw = np.zeros((n, 4), np.uint32)
    ln = np.ones(n, np.uint8)
    with pytest.raises(core.SxError) as e:
        dev.check_pairs_noalloc((tw, ln), (tw, ln))
    assert e.value.code == core.SX_EINVAL and "maximum" in str(e.value)
    z = np.zeros((0, 4), np.uint32), np.zeros(0, np.uint8)
    assert len(dev.check_pairs_noalloc(z, z)) == 0
    # the largest allowed call still matches
    t, c = _pairs(3, core.NOALLOC_MAX_PAIRS)
    assert np.array_equal(dev.check_pairs_noalloc(t, c, ALL, dead=0x82), dev.check_pairs(t, c, ALL, dead=0x82))
