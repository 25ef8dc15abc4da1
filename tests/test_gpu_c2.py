"""The full length-2 enumeration (7,633,612 targets x 2,863 candidates) in Go
semantics at --dead-flags 0x00 and 0xFF, against the oracle.

The fixture tests/golden/c2_strata.npz (make_c2_strata.py) holds the
oracle's rules, pure Go semantics (no MidCheck) with complete sweeps, for
64,009 targets: every target that reads three or more extra registers on
its own and 61,440 drawn uniformly.  The device must emit exactly those
rules, given the targets as a list and inside the full run; the full runs
are also checked for their size-independent properties (counts, every rule
of a sample re-verified, bytes saved, FlagDiff dead flags).
"""
import os

import numpy as np
import pytest

import oracle as o
from sx6502 import core, isa

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
N_L2 = 7_633_612
QC_PAIRS = N_L2 * 2863  # 21,855,031,156 (README.md:132)


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def strata():
    return np.load(os.path.join(GOLD, "c2_strata.npz"), allow_pickle=False)


@pytest.fixture(scope="module")
def l2(dev):
    return dev.pruned_targets()


def _compact(rules, pos):
    """(target position, rep word, pass, dead, bytes, cycles) rows, sorted."""
    rows = np.stack([pos, rules["rep"][:, 0].astype(np.int64), rules["pass_"].astype(np.int64),
                     rules["dead_flags"].astype(np.int64), rules["bytes_saved"].astype(np.int64),
                     rules["cycles_saved"].astype(np.int64)], 1)
    return rows[np.lexsort(rows.T[::-1])]


def _gold(strata, dead):
    p = f"r{dead:02X}_"
    rows = np.stack([strata[p + "t"].astype(np.int64), strata[p + "rep"].astype(np.int64),
                     strata[p + "pass"].astype(np.int64), strata[p + "dead"].astype(np.int64),
                     strata[p + "bytes"].astype(np.int64), strata[p + "cycles"].astype(np.int64)], 1)
    return rows[np.lexsort(rows.T[::-1])]


def _diff(a, b, targets):
    sa, sb = {tuple(r) for r in a.tolist()}, {tuple(r) for r in b.tolist()}
    show = lambda r: f"{isa.disasm_seq(targets[r[0]].tolist())} -> {o.disasm(r[1])} pass {r[2]}"  # noqa: E731
    return [("device only", show(r)) for r in sorted(sa - sb)[:5]] + \
        [("oracle only", show(r)) for r in sorted(sb - sa)[:5]]


@pytest.mark.parametrize("dead", [0x00, 0xFF])
def test_strata_as_a_target_list(dev, strata, l2, dead):
    print(f"c2 strata list, dead 0x{dead:02X}", flush=True)
    targets = l2[strata["index"]]
    assert np.array_equal(targets, o.enumerate_pruned(2)[strata["index"]])
    got, st = dev.enumerate(targets, dead=dead)
    a, b = _compact(got, got["target_index"].astype(np.int64)), _gold(strata, dead)
    assert np.array_equal(a, b), _diff(a, b, targets)
    assert st["targets"] == len(targets)


@pytest.mark.parametrize("dead,n_rules", [(0x00, 1_043_372), (0xFF, 4_418_625)])
def test_full_run(dev, strata, l2, dead, n_rules):
    print(f"c2 full run, dead 0x{dead:02X}", flush=True)
    got, st = dev.enumerate(None, dead=dead)
    assert st["targets"] == N_L2 and st["qc_pairs"] == QC_PAIRS
    assert len(got) == st["rules"] == n_rules
    ti = got["target_index"].astype(np.int64)
    assert (np.diff(ti * 2 + got["pass_"]) > 0).all()  # one rule per (target, pass), in order
    assert (got["bytes_saved"] > 0).all() and (got["src_len"] == 2).all() and (got["rep_len"] == 1).all()
    assert np.array_equal(got["src"][:, :2], l2[ti])
    p2 = got["pass_"] == 2
    assert (got["dead_flags"][~p2] == 0).all()
    if dead:
        assert p2.any() and (got["dead_flags"][p2] != 0).all() and not (got["dead_flags"][p2] & 0x30).any()
    else:
        assert not p2.any()
    # the full run's rules at the stratified targets are the oracle's
    pos = np.searchsorted(strata["index"], ti)
    hit = (pos < len(strata["index"])) & (strata["index"][np.minimum(pos, len(strata["index"]) - 1)] == ti)
    a, b = _compact(got[hit], pos[hit]), _gold(strata, dead)
    assert np.array_equal(a, b), _diff(a, b, l2[strata["index"]])
    # a sample re-verified pair by pair under its pass's mask
    pick = got[:: max(1, len(got) // 6000)]
    for p in (1, 2):
        m = pick["pass_"] == p
        if not m.any():
            continue
        ts = (pick["src"][m], pick["src_len"][m])
        cs = (pick["rep"][m], pick["rep_len"][m])
        ok = dev.check_pairs(ts, cs, core.STAGE_QC | core.STAGE_EXH, 0 if p == 1 else dead)
        assert ((ok & core.STAGE_EXH) != 0).all() and ((ok & core.STAGE_QC) != 0).all()
        if p == 2:  # the flag-relaxed pass takes candidates that fail the exact QuickCheck
            exact = dev.check_pairs(ts, cs, core.STAGE_QC, 0)
            assert not (exact & core.STAGE_QC).any()
            fd = dev.flag_diff([t[:2].tolist() for t in ts[0]], [c[:1].tolist() for c in cs[0]])
            assert np.array_equal(fd, pick["dead_flags"][m])


def test_disjoint_ranges_concatenate_to_the_whole_run(dev):
    """Sharding by target range (first_target / end_target, one range per
    process): three ranges with boundaries inside segments, concatenated,
    are the whole run's rules, record for record, at 0xFF."""
    print("c2 ranges, dead 0xFF", flush=True)
    whole, st = dev.enumerate(None, dead=0xFF)
    cuts = [0, 1_000_003, 4_200_000, N_L2]
    parts = []
    for a, b in zip(cuts[:-1], cuts[1:]):
        r, s = dev.enumerate(None, dead=0xFF, first=a, end=b)
        assert s["targets"] == b - a
        ti = r["target_index"].astype(np.int64)
        assert ((ti >= a) & (ti < b)).all()
        parts.append(r)
    joined = np.concatenate(parts)
    assert joined.tobytes() == whole.tobytes()
