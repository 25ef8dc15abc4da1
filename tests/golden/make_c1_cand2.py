"""The oracle's rules for the 4,096 sampled length-2 targets against every
length-1 and length-2 candidate (CPU only).

    python tests/golden/make_c1_cand2.py [threads]

`6502opt enumerate --max-target 3` runs every target length 2..3 with
candidates of length 1..2 (pkg/search/search.go:26-50), so its length-2
targets also meet the 8,196,769 two-instruction candidates when no single
instruction replaces them.  Go semantics (first hit in Go order, shortest
candidates first, Go byte sizes, no MidCheck), complete sweeps (oracle mode
2), at --dead-flags 0x00 and 0xFF.

Writes c1_cand2.npz with the rule arrays r00 and rFF (RULE_DTYPE records;
target_index = position in c1_targets.npy) and c1_cand2_counts.json.
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as o  # noqa: E402


def main(nthreads: int = 8):
    targets = np.load(os.path.join(HERE, "c1_targets.npy"))
    out, counts = {}, {}
    for dead in (0x00, 0xFF):
        t0 = time.time()
        rules, st = o.search(targets, max_cand_len=2, dead=dead, exh_mode=o.EXH_COMPLETE, use_mid=0,
                             nthreads=nthreads)
        out[f"r{dead:02X}"] = rules
        counts[f"dead{dead:02X}"] = {"rules": int(len(rules)), "len2": int((rules["rep_len"] == 2).sum()),
                                     "checked": st.checked, "qc_hits": st.qc_hits, "exh_calls": st.exh_calls,
                                     "seconds": round(time.time() - t0, 1)}
        print(counts[f"dead{dead:02X}"], flush=True)
    np.savez_compressed(os.path.join(HERE, "c1_cand2.npz"), **out)
    with open(os.path.join(HERE, "c1_cand2_counts.json"), "w") as f:
        json.dump(counts, f, indent=1)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
