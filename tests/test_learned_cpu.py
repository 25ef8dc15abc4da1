"""Host-side pieces of the learned STOKE proposal (sx6502/learned.py): the
target generator, tokenisation and the training labels."""
import os

import numpy as np
import pytest

from sx6502 import core, isa, learned

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
W = isa.word
OP = isa.OP


def test_random_targets_shape_and_stack_rule():
    ts = learned.random_targets(500, seed=7)
    assert len(ts) == 500
    assert all(4 <= len(t) <= 8 for t in ts)
    assert all(learned.stack_safe(t) for t in ts)
    assert ts == learned.random_targets(500, seed=7)
    assert ts != learned.random_targets(500, seed=8)
    lens = np.bincount([len(t) for t in ts], minlength=9)[4:]
    assert lens.min() > 60  # lengths 4..8 roughly uniform


def test_training_targets_do_not_overlap_the_measured_targets():
    t = np.load(os.path.join(GOLD, "c4_targets.npy"))
    ln = np.load(os.path.join(GOLD, "c4_lens.npy"))
    measured = {tuple(t[i, :ln[i]].tolist()) for i in range(len(t))}
    train = learned.random_targets(20000, seed=1000)
    assert not measured & {tuple(s) for s in train}


def test_stack_safe():
    assert learned.stack_safe([W(OP["PHA"]), W(OP["PLA"])])
    assert not learned.stack_safe([W(OP["PLA"])])
    assert not learned.stack_safe([W(OP["PHA"]), W(OP["PLA"]), W(OP["PLP"])])


def test_tokens_pad_and_index():
    seqs = [[W(OP["TXA"]), W(OP["LDA_IMM"], 0x12)], []]
    tok = learned.tokens(seqs)
    assert tok.shape == (2, 8) and tok.dtype == np.uint16
    assert tok[0, 0] == learned.INDEX[W(OP["TXA"])]
    assert learned.TABLE[tok[0, 1]] == W(OP["LDA_IMM"], 0x12)
    assert (tok[0, 2:] == core.PAD_TOKEN).all() and (tok[1] == core.PAD_TOKEN).all()
    with pytest.raises(ValueError):
        learned.tokens([[W(OP["TXA"])] * 9])


def test_labels_point_at_context_positions():
    ldx5, txa, lda9, lda7 = W(OP["LDX_IMM"], 5), W(OP["TXA"]), W(OP["LDA_IMM"], 9), W(OP["LDA_IMM"], 7)
    src = [ldx5, txa, ldx5, lda9]
    ptr = core.MLP_CLASSES
    # drops TXA and one LDX #5: pointers to the first unused positions
    ctx, lab = learned.examples([(src, [ldx5, lda9])], "dropped")
    assert lab.tolist() == [ptr + 0, ptr + 1]
    assert (ctx == learned.tokens([src])).all()
    # kept instructions point at their positions; a new instruction is its index
    _, lab = learned.examples([(src, [ldx5, lda7])], "rewrite")
    assert lab.tolist() == [ptr + 0, learned.INDEX[lda7]]
    _, lab = learned.examples([(src, [ldx5, lda9])], "both")
    assert lab.tolist() == [ptr + 0, ptr + 1, ptr + 2, ptr + 3]
    _, lab = learned.examples([(src, [ldx5, lda9])], "dropped", repeat=3)
    assert len(lab) == 6
    ctx, lab = learned.examples([], "dropped")
    assert ctx.shape == (0, 8) and lab.shape == (0,)
    with pytest.raises(ValueError):
        learned.examples([(src, [])], "nonsense")
