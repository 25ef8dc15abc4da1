"""The oracle's length-3 candidate scan (scan_len3 in oracle/sx_oracle.c)
against a brute force written from worker.go:119-245 with the oracle's
per-pair predicates.

For targets of four one-byte instructions (4 bytes) the byte filter admits
only candidates of three one-byte instructions, so the brute force can walk
all 47^3 of them in Go order and apply the pruner, QuickCheck,
ExhaustiveCheck (complete) and, in the flag-relaxed pass, FlagDiff != 0 one
pair at a time, with no prefix sharing and no early exit on vector 0.  The
first hit (or none) must equal the search's, for targets whose first hit is
of length 3 and for ones that have none.
"""
import numpy as np
import pytest

import oracle as o

IMPLIED = [int(w) for w in o.instruction_table(0)[:47]]  # Go order: the one-byte instructions first


def brute_first(t, dead, pass_):
    for a in IMPLIED:
        for b in IMPLIED:
            for c in IMPLIED:
                cand = [a, b, c]
                if o.should_prune(cand):
                    continue
                if pass_ == 1:
                    if not o.quick_check(t, cand, 0) or not o.exhaustive_check(t, cand, 0, mode=2):
                        continue
                else:
                    if o.quick_check(t, cand, 0) or not o.quick_check(t, cand, dead):
                        continue
                    if not o.exhaustive_check(t, cand, dead, mode=2) or o.flag_diff(t, cand) == 0:
                        continue
                return cand
    return None


W = o.word
TARGETS = [
    [W(o.LDX_M), W(o.LDY_M), W(o.SBC_M), W(o.TAY)],   # -> LDX M : SBC M : TAY
    [W(o.CLV), W(o.STY_M), W(o.BIT_M), W(o.ROL_A)],   # -> STY M : BIT M : ROL A
    [W(o.DEX), W(o.CPY_M), W(o.TYA), W(o.LDX_M)],     # -> TYA : CMP M : LDX M
    [W(o.CLC), W(o.CLV), W(o.LSR_A), W(o.INX)],       # -> LSR A : INX : CLV
    [W(o.TXA), W(o.INY), W(o.ASL_M), W(o.PHA)],       # no shorter form
    [W(o.SEC), W(o.ROR_A), W(o.TAX), W(o.INC_M)],     # no shorter form
]


@pytest.mark.parametrize("dead,idx", [(0x00, i) for i in range(len(TARGETS))] + [(0x01, 2), (0x01, 3), (0x01, 5)])
def test_scan_len3_matches_the_brute_force(dead, idx):
    t = TARGETS[idx]
    short, _ = o.search(np.array([t], np.uint32), max_cand_len=2, dead=dead, exh_mode=o.EXH_COMPLETE, nthreads=1)
    rules, st = o.search(np.array([t], np.uint32), max_cand_len=3, dead=dead, exh_mode=o.EXH_COMPLETE, nthreads=1)
    for pass_ in ((1, 2) if dead else (1,)):
        s = short[short["pass_"] == pass_]
        r = rules[rules["pass_"] == pass_]
        if len(s):  # a first hit of length 1 or 2 is kept as is
            assert len(r) == 1 and r[0].tobytes() == s[0].tobytes()
            continue
        want = brute_first(t, dead, pass_)
        got = r["rep"][0][: r["rep_len"][0]].tolist() if len(r) else None
        assert got == want, (pass_, got, want)
    # every candidate of lengths 1..3 counted once per pass (worker.go:126)
    n = 2863
    assert st.checked >= n + n * n
