"""Batched length-4 targets against candidate pools of length 1..3 (CPU).

    python tests/golden/make_c4_l3.py [threads]

The oracle's first-hit search (worker.go:119-245: candidate lengths 1, 2, 3
in turn, Go order within a length; the exact pass, then with dead flags the
flag-relaxed pass) with complete sweeps, over 32 length-4 targets: 16 whose
effects a shorter sequence can reproduce (redundant loads, stores and flag
operations) and 16 pruned targets drawn with numpy's PCG64 (seed 4) from the
2,863 instructions.  Length-3 candidates (2,863^3 of them) are enumerated on
the fly; a target without a shorter form costs a full scan (minutes of CPU).
Writes c4_l3.npz: targets, and one RULE_DTYPE array per run (dead 0x00 over
all 32 targets, dead 0x01 over the first 16).
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as o  # noqa: E402


def constructed():
    W = o.word
    return [
        [W(o.LDA_IMM, 1), W(o.LDX_IMM, 2), W(o.LDY_IMM, 3), W(o.LDA_IMM, 4)],
        [W(o.CLC), W(o.LDA_IMM, 5), W(o.ADC_IMM, 3), W(o.TAX)],
        [W(o.TAX), W(o.TXA), W(o.TAY), W(o.TYA)],
        [W(o.LDA_M), W(o.STA_M), W(o.LDX_IMM, 7), W(o.INX)],
        [W(o.SEC), W(o.CLC), W(o.LDA_IMM, 0x80), W(o.ROL_A)],
        [W(o.LDX_IMM, 0), W(o.INX), W(o.INX), W(o.DEX)],
        [W(o.PHA), W(o.PLA), W(o.TAX), W(o.CLC)],
        [W(o.LDA_IMM, 0x0F), W(o.AND_IMM, 0xF0), W(o.ORA_IMM, 0x33), W(o.TAY)],
        [W(o.INC_M), W(o.DEC_M), W(o.LDA_M), W(o.CMP_IMM, 4)],
        [W(o.LDY_IMM, 9), W(o.TYA), W(o.TAX), W(o.LDY_IMM, 1)],
        [W(o.SEC), W(o.LDA_IMM, 0x10), W(o.SBC_IMM, 0x01), W(o.STA_M)],
        [W(o.TXA), W(o.EOR_IMM, 0xFF), W(o.EOR_IMM, 0xFF), W(o.TAX)],
        [W(o.CLV), W(o.CLC), W(o.LDA_IMM, 2), W(o.ASL_A)],
        [W(o.LDX_M), W(o.STX_M), W(o.LDY_M), W(o.CPY_IMM, 0)],
        [W(o.TSX), W(o.INX), W(o.DEX), W(o.TXS)],
        [W(o.LDA_IMM, 0), W(o.ORA_M), W(o.AND_M), W(o.STA_M)],
    ]


def random_targets(n, seed=4):
    rng = np.random.default_rng(seed)
    tab = o.instruction_table(0)
    out = []
    while len(out) < n:
        t = [int(tab[i]) for i in rng.integers(0, len(tab), 4)]
        if not o.should_prune(t):
            out.append(t)
    return out


def main(nthreads: int = 8):
    targets = np.array(constructed() + random_targets(16), np.uint32)
    out = {"targets": targets}
    for name, sel, dead in (("dead00", slice(None), 0x00), ("dead01", slice(0, 16), 0x01)):
        t0 = time.time()
        rules, st = o.search(targets[sel], max_cand_len=3, dead=dead, exh_mode=o.EXH_COMPLETE, nthreads=nthreads)
        out[name] = rules
        print(name, len(rules), "rules,", st.checked, "candidates checked, %.0f s" % (time.time() - t0), flush=True)
        for r in rules:
            dis = lambda ws, n: " : ".join(o.disasm(int(w)) for w in ws[:n])  # noqa: E731
            print("  %2d pass %d  %s -> %s  dead 0x%02X" % (r["target_index"], r["pass_"], dis(r["src"], r["src_len"]),
                                                          dis(r["rep"], r["rep_len"]), r["dead_flags"]), flush=True)
    np.savez_compressed(os.path.join(HERE, "c4_l3.npz"), **out)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
