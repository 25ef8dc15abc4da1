"""STOKE on the device against the oracle's chain restatement.

Trajectory parity: the device chains and the oracle's chains, started at the
same targets with the same Philox key and chain indices, end in the same
state (sequence, cost, accepted count) after thousands of steps, with the
uniform and with a non-uniform instruction distribution.  Every rewrite the
device reports passes the exhaustive check and saves bytes.
"""
import os

import numpy as np
import pytest

import oracle as o
from sx6502 import core, isa

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
M32 = 1 << 32


def c4_targets(k=None):
    t = np.load(os.path.join(GOLD, "c4_targets.npy"))
    ln = np.load(os.path.join(GOLD, "c4_lens.npy"))
    out = [t[i, : ln[i]].tolist() for i in range(len(t))]
    return out[:k] if k else out


def random_weights(n, seed):
    rng = np.random.default_rng(seed)
    w = np.zeros((n, 2863), np.uint64)
    for i in range(n):
        p = 0.9 * rng.dirichlet(np.full(2863, 0.3)) + 0.1 / 2863
        w[i] = np.floor(p * M32).astype(np.uint64)
        w[i, np.argmax(w[i])] += M32 - int(w[i].sum())
    return w.astype(np.uint32)


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("learned", [False, True])
def test_trajectories_match_oracle(dev, learned):
    targets = c4_targets(6)
    chains, steps, seed, decay = 4, 3000, 77, 0.999
    w = random_weights(len(targets), 5) if learned else None
    res, ch = dev.stoke(targets, chains=chains, steps=steps, steps_per_launch=1000, decay=decay, seed=seed,
                        weights=w, stop_when_verified=False, return_chains=True)
    for ti, t in enumerate(targets):
        for j in range(chains):
            c = ti * chains + j
            oc = o.stoke_chain(t)
            o.stoke_steps(t, oc, steps, chain_id=c, seed=seed, decay=decay,
                          weights=None if w is None else w[ti])
            g = ch[c]
            assert g["step"] == steps
            assert list(g["seq"][: g["len"]]) == list(oc.seq[: oc.len]), (ti, j)
            assert g["cost10"] == oc.cost10 and g["accepted"] == oc.accepted, (ti, j)


def test_reported_rewrites_are_sound(dev):
    targets = c4_targets(48)
    res = dev.stoke(targets, chains=16, steps=20000, seed=3)
    assert res["verified"].sum() > 0
    for t, r in zip(targets, res):
        if not r["verified"]:
            continue
        rep = r["rep"][: r["rep_len"]].tolist()
        assert r["bytes_saved"] == isa.seq_byte_size(t) - isa.seq_byte_size(rep) > 0
        assert o.quick_check(t, rep) and o.mid_check(t, rep)
        assert o.exhaustive_check(t, rep, 0, mode=0)   # necessary: the complete sweep implies it
