"""Batched length-4 targets against candidate pools of length 1..3
(sx_enumerate with target_len 4, max_cand_len 3).

Parity: the 32 targets of tests/golden/c4_l3.npz (make_c4_l3.py: the
oracle's first-hit search over lengths 1, 2, 3 with complete sweeps, the
length-3 candidates enumerated on the fly) give exactly the oracle's rules
at --dead-flags 0x00 and 0x01.  At a larger batch, size-independent
properties: the rules of lengths 1 and 2 are those of the same batch with
max_cand_len 2 (the first hit goes by candidate length first), length-3
rules appear only for entries that run leaves open, and every rule passes
the complete ExhaustiveCheck again.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("target_index", "pass_", "src", "src_len", "rep", "rep_len", "dead_flags", "bytes_saved", "cycles_saved")


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def gold():
    return np.load(os.path.join(GOLD, "c4_l3.npz"), allow_pickle=False)


@pytest.mark.parametrize("name,n,dead", [("dead00", 32, 0x00), ("dead01", 16, 0x01)])
def test_matches_the_oracle(dev, gold, name, n, dead):
    print(f"length-4 targets x length 1..3, {name}", flush=True)
    targets = gold["targets"][:n]
    got, st = dev.enumerate(targets, target_len=4, max_cand_len=3, dead=dead)
    want = gold[name]
    assert len(got) == len(want), (len(got), len(want))
    for f in FIELDS:
        assert np.array_equal(got[f], want[f]), f
    assert (want["rep_len"] == 3).any()  # the fixture reaches the length-3 pool


def test_batch_properties(dev):
    rng = np.random.default_rng(9)
    import oracle as o
    tab = o.instruction_table(0)
    targets = []
    while len(targets) < 2048:
        t = [int(tab[i]) for i in rng.integers(0, len(tab), 4)]
        if not o.should_prune(t):
            targets.append(t)
    targets = np.array(targets, np.uint32)
    print("length-4 batch of 2048, dead 0x01", flush=True)
    r3, _ = dev.enumerate(targets, target_len=4, max_cand_len=3, dead=0x01)
    r2, _ = dev.enumerate(targets, target_len=4, max_cand_len=2, dead=0x01)
    short = r3[r3["rep_len"] <= 2]
    for f in FIELDS:
        assert np.array_equal(short[f], r2[f]), f
    closed = set(zip(r2["target_index"].tolist(), r2["pass_"].tolist()))
    long3 = r3[r3["rep_len"] == 3]
    assert not closed & set(zip(long3["target_index"].tolist(), long3["pass_"].tolist()))
    assert (r3["bytes_saved"] > 0).all()
    for p in (1, 2):
        m = r3["pass_"] == p
        if not m.any():
            continue
        ts = (r3["src"][m], r3["src_len"][m])
        cs = (r3["rep"][m], r3["rep_len"][m])
        ok = dev.check_pairs(ts, cs, core.STAGE_QC | core.STAGE_EXH, 0 if p == 1 else 0x01)
        assert ((ok & core.STAGE_EXH) != 0).all() and ((ok & core.STAGE_QC) != 0).all()
