"""The reduced sweep against the complete sweep, on the published full
length-2 run's CPU-fallback pairs (CPU only; nothing here calls the product).

The CUDA tool sends every QuickCheck + MidCheck survivor that reads three or
more extra registers to a CPU sweep over 32 representative values per
register (cuda/6502_search_v2.cu:650,671-736,945-974; Go's
exhaustiveReducedSweep, pkg/search/verifier.go:377-415).  The fixture
tests/golden/c2_cuda_fallback_k3.npz (make_fallback_k3.py) lists those pairs
and their verdicts.  These tests pin:

* the pair list: the oracle's pipeline alone yields the published fallback
  count, 4,902 (README.md:135);
* the gap: the reduced sweep passes 4,118 of them and the complete sweep
  1,026, and 4,118 - 1,026 = 3,092 = 1,078,897 (README.md:136) - 1,075,805
  (the device's rule count for that run, test_gpu_core.py);
* every rejection by a domain state on which the scalar executor shows the
  two sequences disagree (a reduced-only pass has such a state with an extra
  register outside the 32 values, so the reduced sweep cannot see it);
* the complete verdicts on a sample by brute force over all 2^33 states;
* the oracle's reduced sweep against a numpy brute force of the 32^3 x 512
  grid with the independent textbook model.
"""
import os

import numpy as np
import pytest

import oracle as o
from textbook import model

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FIX = np.load(os.path.join(GOLD, "c2_cuda_fallback_k3.npz"), allow_pickle=False)

PUBLISHED_FALLBACK = 4_902     # README.md:135
PUBLISHED_RULES = 1_078_897    # README.md:136
DEVICE_RULES = 1_075_805       # the device's complete-sweep count of that run (test_gpu_core.py)
# verifier.go:378-383
REP = [0x00, 0x01, 0x02, 0x0F, 0x10, 0x1F, 0x20, 0x3F, 0x40, 0x55, 0x7E, 0x7F, 0x80, 0x81, 0xAA, 0xBF,
       0xC0, 0xD5, 0xE0, 0xEF, 0xF0, 0xF7, 0xFE, 0xFF, 0x03, 0x07, 0x11, 0x33, 0x77, 0xBB, 0xDD, 0xEE]
A, X, Y, P, S, M, S0, S1 = range(8)
# extra registers in Go's order X, Y, M, S, S0, S1 (verifier.go:292-310) with their read bits
EXTRAS = [(X, 4), (Y, 8), (M, 32), (S, 16), (S0, 64), (S1, 128)]


def _seqs(i):
    p = FIX["pairs"][i]
    return [int(p[0]), int(p[1])], [int(p[2])]


def _extras(t, c):
    reads = 0
    for w in t + c:
        reads |= o.op_reads(w & 0xFFFF)
    return [b for b, bit in EXTRAS if reads & bit]


def _disagree(st, t, c):
    a = o.exec_seq(list(st), t)
    b = o.exec_seq(list(st), c)
    return any(x != y for k, (x, y) in enumerate(zip(a, b)) if k != P) or ((a[P] ^ b[P]) & 0xCF) != 0


def _in_domain(st, extras):
    # A any, P = carry only, S = 0xFD unless swept, unread bytes 0 (verifier.go:262-415)
    if st[P] > 1:
        return False
    for b in (X, Y, M, S0, S1):
        if b not in extras and st[b] != 0:
            return False
    return S in extras or st[S] == 0xFD


def test_fallback_pairs_are_the_published_count():
    pairs = o.cuda_fallback_pairs(3, nthreads=8)
    assert len(pairs) == PUBLISHED_FALLBACK
    assert np.array_equal(pairs, FIX["pairs"])
    assert all(len(_extras(*_seqs(i))) >= 3 for i in range(0, len(pairs), 97))


def test_reduced_minus_complete_is_the_published_gap():
    red, com = FIX["reduced"], FIX["complete"]
    assert not (com & ~red).any()  # the reduced grid lies inside the domain
    assert int(red.sum()) == 4_118 and int(com.sum()) == 1_026
    assert int(red.sum()) - int(com.sum()) == PUBLISHED_RULES - DEVICE_RULES == 3_092


def test_every_rejection_has_a_real_counterexample():
    red, com, cex, rcex = FIX["reduced"], FIX["complete"], FIX["cex"], FIX["rcex"]
    outside = 0
    for i in np.nonzero(~com)[0]:
        t, c = _seqs(i)
        ex = _extras(t, c)
        st = cex[i].tolist()
        assert _in_domain(st, ex), (i, st)
        assert _disagree(st, t, c), (i, st)
        if red[i]:
            # the reduced sweep passed it: some swept register is off its 32 values
            assert any(st[b] not in REP for b in ex), (i, st)
            outside += 1
        else:
            st = rcex[i].tolist()
            assert _in_domain(st, ex) and all(st[b] in REP for b in ex) and _disagree(st, t, c), (i, st)
    assert outside == 3_092


def test_verdicts_recomputed_on_a_sample():
    from concurrent.futures import ThreadPoolExecutor  # the oracle's calls release the GIL

    def both(i):
        t, c = _seqs(i)
        return o.exhaustive_check(t, c, 0, o.EXH_GO), o.exhaustive_check(t, c, 0, o.EXH_COMPLETE)
    idx = list(range(0, PUBLISHED_FALLBACK, 61))
    with ThreadPoolExecutor(8) as ex:
        got = list(ex.map(both, idx))
    assert [g[0] for g in got] == FIX["reduced"][idx].tolist()
    assert [g[1] for g in got] == FIX["complete"][idx].tolist()


def test_complete_verdicts_by_brute_force():
    """The fixture's brute-force sample (the whole 2^33-state domain) agrees
    with the complete verdicts; one pass and one reduced-only pass are
    brute-forced again here."""
    idx, brute = FIX["brute_idx"], FIX["brute"]
    assert len(idx) >= 48 and brute.sum() >= 40
    assert np.array_equal(brute, FIX["complete"][idx])
    for i in (idx[0], idx[-1]):
        t, c = _seqs(i)
        ok, cex = o.exhaustive_brute(t, c, 0, nthreads=8)
        assert ok == bool(FIX["complete"][i])
        if not ok:
            assert _disagree(cex.tolist(), t, c)


def _reduced_sweep_numpy(t, c):
    """verifier.go:377-415 with the textbook model: every combination of the
    32 values for each extra register, A over 0..255, carry 0/1."""
    ex = _extras(t, c)
    grid = np.stack(np.meshgrid(*([np.array(REP)] * (len(ex) - 1)), indexing="ij"), -1).reshape(-1, len(ex) - 1)
    ac = np.stack(np.meshgrid(np.arange(256), np.arange(2), indexing="ij"), -1).reshape(-1, 2)
    for v0 in REP:  # one value of the first extra register per chunk
        st = np.zeros((len(grid) * len(ac), 8), np.int64)
        st[:, S] = 0xFD
        st[:, ex[0]] = v0
        for k, b in enumerate(ex[1:]):
            st[:, b] = np.repeat(grid[:, k], len(ac))
        st[:, A] = np.tile(ac[:, 0], len(grid))
        st[:, P] = np.tile(ac[:, 1], len(grid))
        ts, cs = st.astype(np.uint8), st.astype(np.uint8)
        for w in t:
            ts = model(w & 0xFFFF, ts, w >> 16)
        for w in c:
            cs = model(w & 0xFFFF, cs, w >> 16)
        d = (ts ^ cs).astype(np.int64)
        d[:, P] &= 0xCF
        if d.any():
            return False
    return True


def test_reduced_sweep_against_textbook_bruteforce():
    red, com = FIX["reduced"], FIX["complete"]
    pick = list(np.nonzero(red & ~com)[0][[0, 2000]]) + list(np.nonzero(com)[0][[500]]) + \
        list(np.nonzero(~red)[0][[300]])
    for i in pick:
        t, c = _seqs(i)
        assert _reduced_sweep_numpy(t, c) == bool(red[i]), (i, [o.disasm(w) for w in t + c])
