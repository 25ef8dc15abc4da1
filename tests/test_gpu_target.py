"""The `target` command on the device (sx_target_search): candidates
enumerated on the fly in Go order, against the oracle's rule sets (C1
fixtures), against the table-driven enumeration (length-3 targets with
length-1/2 candidates) and against known answers with length-3 candidates."""
import os

import numpy as np
import pytest

import oracle as o
from sx6502 import core, isa

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
W, OP = isa.word, isa.OP


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


def fixture_rules(name):
    rules = np.load(os.path.join(GOLD, name + ".npy"))
    out = {}
    for r in rules:
        out[(int(r["target_index"]), int(r["pass_"]))] = [int(w) for w in r["rep"][: r["rep_len"]]]
    return out


@pytest.mark.parametrize("name,dead", [("c1_go_dead00", 0x00), ("c1_go_deadFF", 0xFF), ("c1_go_dead01", 0x01)])
def test_c1_rules_match_the_oracle(dev, name, dead):
    targets = np.load(os.path.join(GOLD, "c1_targets.npy"))
    want = fixture_rules(name)
    # every target with a rule, plus as many without
    with_rule = sorted({ti for ti, _ in want})
    without = [i for i in range(len(targets)) if i not in set(with_rule)][: len(with_rule)]
    print(f"{name}: {len(with_rule)} targets with rules, {len(without)} without", flush=True)
    for ti in with_rule + without:
        res, scanned = dev.target_search(targets[ti].tolist(), max_cand_len=1, dead=dead)
        for p in (1, 2):
            got = res[p - 1]
            exp = want.get((ti, p))
            assert (got["rep"] if got else None) == exp, (ti, p, got, exp)
        assert scanned == 2863


def test_length3_targets_match_the_table_driven_enumeration(dev):
    targets = np.load(os.path.join(GOLD, "c3_targets.npy"))[:256]
    rules, _ = dev.enumerate(targets, target_len=3, max_cand_len=2, dead=0x01)
    want = {(int(r["target_index"]), int(r["pass_"])): [int(w) for w in r["rep"][: r["rep_len"]]] for r in rules}
    for ti in range(len(targets)):
        res, _ = dev.target_search(targets[ti].tolist(), max_cand_len=2, dead=0x01)
        for p in (1, 2):
            got = res[p - 1]
            assert (got["rep"] if got else None) == want.get((ti, p)), (ti, p)


def test_known_answers_with_longer_candidates(dev):
    lda, ldx, ldy = (lambda v: W(OP["LDA_IMM"], v)), (lambda v: W(OP["LDX_IMM"], v)), (lambda v: W(OP["LDY_IMM"], v))
    # one redundant load in four: the first length-3 candidate in Go order
    t = [lda(1), ldx(2), lda(3), ldy(4)]
    res, scanned = dev.target_search(t, max_cand_len=3)
    assert res[0]["rep"] == [lda(3), ldx(2), ldy(4)]
    assert res[0]["bytes_saved"] == 2 and res[0]["cycles_saved"] == 2 and res[1] is None
    assert o.quick_check(t, res[0]["rep"]) and o.exhaustive_check(t, res[0]["rep"], 0, 1)
    # every length-1 and length-2 candidate, then length-3 ones up to the launch that held the hit
    assert 2863 + 2863 ** 2 < scanned < 2863 + 2863 ** 2 + 2863 ** 3
    assert dev.target_search(t, max_cand_len=2)[0] == [None, None]
    # TXA : TAX -> TXA (README example)
    res, _ = dev.target_search([W(OP["TXA"]), W(OP["TAX"])], max_cand_len=4)
    assert res[0]["rep"] == [W(OP["TXA"])]
    # CLC : LDA #5 only shortens when carry is dead
    t = [W(OP["CLC"]), lda(5)]
    res, _ = dev.target_search(t, max_cand_len=2, dead=0x01)
    assert res[0] is None
    assert res[1]["rep"] == [lda(5)] and res[1]["dead_flags"] == 0x01 and res[1]["bytes_saved"] == 1
    # nothing shorter exists for a single instruction
    assert dev.target_search([lda(7)], max_cand_len=4)[0] == [None, None]


def test_rejects_bad_arguments(dev):
    with pytest.raises(core.SxError):
        dev.target_search([], max_cand_len=1)
    with pytest.raises(core.SxError):
        dev.target_search([W(OP["TXA"])] * 9, max_cand_len=1)
    with pytest.raises(core.SxError):
        dev.target_search([W(OP["TXA"])], max_cand_len=5)
    with pytest.raises(core.SxError):
        dev.target_search([58], max_cand_len=1)
