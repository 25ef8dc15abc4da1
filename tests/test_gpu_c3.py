"""Length-3 targets against all length-1 and length-2 candidates (the
batched configuration), Go semantics, --dead-flags 0x01.

The fixture holds the oracle's rules for the first 512 targets with complete
sweeps (Go's semantics with its reduced sweep for >= 3 extra registers
replaced by a complete one, make_golden_c3.py); the device's rule set must
be identical.
"""
import os

import numpy as np
import pytest

from sx6502 import core, isa

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("src", "rep", "target_index", "src_len", "rep_len", "dead_flags", "pass_", "bytes_saved",
          "cycles_saved")


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


def test_c3_subset_against_oracle(dev):
    gold = np.load(os.path.join(GOLD, "c3_dead01_512.npy"))
    targets = np.load(os.path.join(GOLD, "c3_targets.npy"))[:512]
    got, st = dev.enumerate(targets, target_len=3, max_cand_len=2, dead=0x01)
    assert st["targets"] == 512
    key = lambda r: (int(r["target_index"]), int(r["pass_"]))  # noqa: E731
    g, d = {key(r): r for r in gold}, {key(r): r for r in got}
    diff = [(k, isa.disasm_seq(targets[k[0]].tolist()),
             isa.disasm_seq(g[k]["rep"][: g[k]["rep_len"]].tolist()) if k in g else None,
             isa.disasm_seq(d[k]["rep"][: d[k]["rep_len"]].tolist()) if k in d else None)
            for k in sorted(set(g) | set(d))
            if k not in g or k not in d or any(not np.array_equal(g[k][f], d[k][f]) for f in FIELDS)]
    assert not diff, diff[:5]
    assert len(gold) == len(got) > 200


def test_c3_full_batch_rules_verify(dev):
    """All 65,536 targets in one call; every emitted rule passes the complete
    exhaustive check again (as verify-jsonl would run it) and saves bytes."""
    targets = np.load(os.path.join(GOLD, "c3_targets.npy"))
    got, st = dev.enumerate(targets, target_len=3, max_cand_len=2, dead=0x01)
    assert st["targets"] == 65536 and st["qc_pairs"] == 65536 * (2863 + 2863 * 2863)
    assert len(got) > 0 and (got["bytes_saved"] > 0).all()
    pick = got[:: max(1, len(got) // 4000)]
    ts = [r["src"][:3].tolist() for r in pick]
    cs = [r["rep"][: r["rep_len"]].tolist() for r in pick]
    for p in (1, 2):
        m = pick["pass_"] == p
        if m.any():
            ok = dev.check_pairs([t for t, k in zip(ts, m) if k], [c for c, k in zip(cs, m) if k],
                                 core.STAGE_QC | core.STAGE_EXH, 0 if p == 1 else 0x01)
            assert ((ok & 4) != 0).all()
            if p == 1:
                assert ((ok & 1) != 0).all()
