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


def test_collect_keeps_the_first_mode_rewrites_and_more(dev):
    """sx_stoke_collect (stoke.Run + Deduplicate): every kept rewrite is sound
    and unique per (target, replacement); a target's earliest kept rewrite is
    the first-verified result of sx_stoke_run on the same chains (the two
    runs' trajectories agree until a collecting chain restarts)."""
    targets = c4_targets(48)
    kw = dict(chains=8, steps=12288, seed=5, steps_per_launch=4096)
    first = dev.stoke(targets, stop_when_verified=False, **kw)
    found, nf = dev.stoke_collect(targets, **kw)
    assert nf == len(found) and nf >= first["verified"].sum() > 0
    keys = {(int(f["target"]), tuple(f["rep"][: f["rep_len"]].tolist())) for f in found}
    assert len(keys) == nf
    for f in found:
        t, rep = targets[f["target"]], f["rep"][: f["rep_len"]].tolist()
        assert f["chain"] // 8 == f["target"] and f["dead_flags"] == 0
        assert f["bytes_saved"] == isa.seq_byte_size(t) - isa.seq_byte_size(rep) > 0
        assert o.quick_check(t, rep) and o.mid_check(t, rep) and o.exhaustive_check(t, rep, 0, mode=2)
    for ti, r in enumerate(first):
        if not r["verified"]:
            continue
        mine = found[found["target"] == ti]
        assert len(mine) and mine[0]["step"] == r["step"] and mine[0]["chain"] == r["chain"], ti
        assert mine[0]["rep"].tolist() == r["rep"].tolist()


def test_collect_restarts_and_reports_flag_diff():
    """With carry dead, CLC : CLC : CLC has several one-byte rewrites (CLC,
    SEC, NOP ...): chains restarted after each verified rewrite find more than
    one, and each carries FlagDiff(target, rewrite) as its dead flags
    (pkg/stoke/search.go:104-110), 0 for CLC itself."""
    with core.Device(0) as d:
        t = [o.word(o.CLC)] * 3
        found, nf = d.stoke_collect([t], chains=16, steps=20000, seed=2, dead=0x01, t0=0.05)
        d1, nf1 = d.stoke_collect([t], chains=16, steps=20000, seed=2, dead=0x01, t0=0.05)
    assert nf >= 2, [isa.disasm_seq(f["rep"][: f["rep_len"]].tolist()) for f in found]
    for f in ("target", "chain", "step", "rep", "rep_len", "dead_flags"):  # deterministic for a seed
        assert nf1 == nf and np.array_equal(found[f], d1[f]), f
    for f in found:
        rep = f["rep"][: f["rep_len"]].tolist()
        assert o.exhaustive_check(t, rep, 0x01, mode=2)
        assert f["dead_flags"] == o.flag_diff(t, rep)
    reps = {isa.disasm_seq(f["rep"][: f["rep_len"]].tolist()) for f in found}
    assert len(reps) == nf
    if "CLC" in reps:
        assert [f["dead_flags"] for f in found if f["rep_len"] == 1 and f["rep"][0] == o.word(o.CLC)] == [0]
