"""The exhaustive sweep's two kernels decide alike.

Jobs of length-2 targets against single instructions whose sub-sweeps vary
at most one 8-bit register and the carry are swept one job per lane, sorted
by (opcode triple, compare mask) (k_exh_group); every other job one chunk of
states per wave (k_exh_run).  SX_EXH_GROUP (read when a handle opens) picks
the pairs per lane of the first, or 0 to send every job to the second.  The
rules must be the same record for record, and equal to the oracle's on the
stratified length-2 fixture.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _open(group):
    old = os.environ.get("SX_EXH_GROUP")
    os.environ["SX_EXH_GROUP"] = str(group)
    try:
        return core.Device(0)
    finally:
        if old is None:
            del os.environ["SX_EXH_GROUP"]
        else:
            os.environ["SX_EXH_GROUP"] = old


@pytest.fixture(scope="module")
def devs():
    d = {g: _open(g) for g in (0, 1, 4, 8)}
    yield d
    for x in d.values():
        x.close()


@pytest.fixture(scope="module")
def strata_targets(devs):
    s = np.load(os.path.join(GOLD, "c2_strata.npz"), allow_pickle=False)
    return devs[0].pruned_targets()[s["index"]], s


def _gold_rows(s, dead):
    p = f"r{dead:02X}_"
    rows = np.stack([s[p + "t"].astype(np.int64), s[p + "rep"].astype(np.int64), s[p + "pass"].astype(np.int64),
                     s[p + "dead"].astype(np.int64)], 1)
    return rows[np.lexsort(rows.T[::-1])]


def _rows(r):
    rows = np.stack([r["target_index"].astype(np.int64), r["rep"][:, 0].astype(np.int64),
                     r["pass_"].astype(np.int64), r["dead_flags"].astype(np.int64)], 1)
    return rows[np.lexsort(rows.T[::-1])]


@pytest.mark.parametrize("dead", [0x00, 0xFF])
def test_kernels_agree_on_the_strata(devs, strata_targets, dead):
    print(f"sweep kernels on the strata, dead 0x{dead:02X}", flush=True)
    targets, s = strata_targets
    out = {}
    for g, d in devs.items():
        r, st = d.enumerate(targets, dead=dead)
        out[g] = r
        if g == 0:
            assert st["exh_group_jobs"] == 0 and st["exh_group_states"] == 0 and st["exh_group_launches"] == 0
        else:
            # most jobs of these targets have small plans (the strata hold every
            # target that reads three or more extra registers)
            assert st["exh_group_jobs"] > 0.5 * st["exh_pairs"], st
            assert 0 < st["exh_group_states"] <= st["exh_states"]
    for g in (1, 4, 8):
        assert out[g].tobytes() == out[0].tobytes(), f"SX_EXH_GROUP={g} differs from the state-parallel sweep"
    assert np.array_equal(_rows(out[4]), _gold_rows(s, dead))


def test_kernels_agree_on_the_full_run(devs):
    print("sweep kernels on the full run, dead 0xFF", flush=True)
    a, sa = devs[0].enumerate(None, dead=0xFF)
    b, sb = devs[4].enumerate(None, dead=0xFF)
    assert a.tobytes() == b.tobytes()
    assert sa["exh_pairs"] == sb["exh_pairs"] and sb["exh_group_jobs"] > 0
