"""Length-2 targets against candidates of length 1 and 2 (the length-2 part
of `--max-target 3`, pkg/search/search.go:26-50), against the oracle.

The fixture tests/golden/c1_cand2.npz (make_c1_cand2.py) holds the oracle's
rules for the 4,096 sampled targets in Go semantics with complete sweeps at
--dead-flags 0x00 and 0xFF; the device's must be identical.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIELDS = ("src", "rep", "target_index", "src_len", "rep_len", "dead_flags", "pass_", "bytes_saved",
          "cycles_saved")


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


@pytest.mark.parametrize("dead", [0x00, 0xFF])
def test_c1_targets_with_two_instruction_candidates(dev, dead):
    print(f"c1 targets x (1 + 2)-instruction candidates, dead 0x{dead:02X}", flush=True)
    gold = np.load(os.path.join(GOLD, "c1_cand2.npz"), allow_pickle=False)[f"r{dead:02X}"]
    targets = np.load(os.path.join(GOLD, "c1_targets.npy"))
    got, st = dev.enumerate(targets, target_len=2, max_cand_len=2, dead=dead)
    assert st["qc_pairs"] == 4096 * (2863 + 2863 * 2863)
    assert len(got) == len(gold) and (gold["rep_len"] == 2).any()
    for f in FIELDS:
        assert np.array_equal(got[f], gold[f]), f


def test_all_hits_with_two_instruction_candidates_counts(dev):
    """The CUDA tool's semantics (every hit, M ops 2 bytes, one masked pass)
    with candidates of length 1 and 2 on 512 of the sampled targets: the
    QuickCheck hits after the byte filter and the pruner, the MidCheck
    passes and the rule set equal the oracle's all-hits search.  The hit
    counter also stays the same from run to run (it once counted pruned
    candidates, depending on which survivors shared a wave)."""
    import oracle as o
    targets = np.load(os.path.join(GOLD, "c1_targets.npy"))[:512]
    want, ost = o.search(targets, max_cand_len=2, hits=1, bytes_model=1, exh_mode=o.EXH_COMPLETE, nthreads=8)
    runs = [dev.enumerate(targets, target_len=2, max_cand_len=2, hits=core.HITS_ALL, bytes_model=1)
            for _ in range(2)]
    for got, st in runs:
        assert st["qc_hits"] == ost.qc_hits and st["mid_pass"] == ost.mid_pass
        key = lambda r: (int(r["target_index"]), tuple(r["rep"].tolist()))  # noqa: E731
        assert sorted(map(key, got)) == sorted(map(key, want))
