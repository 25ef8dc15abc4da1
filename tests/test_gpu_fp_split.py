"""The joined QuickCheck + MidCheck with split fingerprints agrees with the
whole-target fingerprints.

With SX_FP_SPLIT=1 (the default) and single-instruction candidates the
targets' outputs on the 8 QuickCheck vectors come first (k_fp_qc), the
QuickCheck class lookup lists the targets whose masked outputs equal some
candidate's, and only those get Mid hashes and enter the join.  With SX_FP_SPLIT=0 every target gets all 32 vectors and
the join reads every target.  Rules and every stage counter must be the same.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
COUNTERS = ("qc_pairs", "qc_raw_hits", "qc_hits", "mid_pass", "exh_pairs", "exh_rounds", "rules")


def _open(split):
    old = os.environ.get("SX_FP_SPLIT")
    os.environ["SX_FP_SPLIT"] = str(split)
    try:
        return core.Device(0)
    finally:
        if old is None:
            del os.environ["SX_FP_SPLIT"]
        else:
            os.environ["SX_FP_SPLIT"] = old


@pytest.fixture(scope="module")
def devs():
    d = {s: _open(s) for s in (0, 1)}
    yield d
    for x in d.values():
        x.close()


def _same(a, b):
    (ra, sa), (rb, sb) = a, b
    assert ra.tobytes() == rb.tobytes()
    for k in COUNTERS:
        assert sa[k] == sb[k], (k, sa[k], sb[k])


@pytest.mark.parametrize("dead", [0x00, 0xFF])
def test_full_length2_run(devs, dead):
    print(f"split fingerprints, full length-2 run, dead 0x{dead:02X}", flush=True)
    _same(devs[0].enumerate(None, dead=dead), devs[1].enumerate(None, dead=dead))


def test_cuda_semantics(devs):
    print("split fingerprints, full length-2 run, all hits", flush=True)
    a = devs[0].enumerate(None, dead=0, hits=core.HITS_ALL, bytes_model=1)
    b = devs[1].enumerate(None, dead=0, hits=core.HITS_ALL, bytes_model=1)
    _same(a, b)
    assert b[1]["qc_hits"] == 20_654_396 and b[1]["mid_pass"] == 5_373_297


@pytest.mark.parametrize("ml", [1, 2])
def test_length3_sample(devs, ml):
    # against single instructions the QuickCheck outputs of length-3 targets
    # go 16 per wave; against 8.2M candidates (max_cand_len 2) the split is
    # not used and both handles take the whole fingerprints
    print(f"split fingerprints, length-3 sample, max_cand_len {ml}", flush=True)
    t = np.load(os.path.join(GOLD, "c3_targets.npy"))[:8192]
    _same(devs[0].enumerate(t, target_len=3, max_cand_len=ml, dead=0x01),
          devs[1].enumerate(t, target_len=3, max_cand_len=ml, dead=0x01))
