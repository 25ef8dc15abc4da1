"""Golden fixtures for the batched length-3 configuration (C3).

    python tests/golden/make_golden_c3.py

Writes c3_targets.npy (65,536 length-3 targets drawn by rejection with
SplitMix64 seed 1, see sampling.py) and the oracle's rules for its first 256
targets against all length-1 and length-2 candidates, Go semantics,
--dead-flags 0x01 (c3_go_dead01_256.npy; Go verifies pairs with >= 3 extra
registers by its 32-value reduced sweep).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, HERE)

import oracle as o  # noqa: E402
import sampling  # noqa: E402


def main():
    tab = o.instruction_table(0)
    t0 = time.time()
    targets = sampling.c3_targets(tab, o.should_prune)
    np.save(os.path.join(HERE, "c3_targets.npy"), targets)
    print("sampled", targets.shape, round(time.time() - t0, 1), flush=True)
    t0 = time.time()
    rules, st = o.search(targets[:256], max_cand_len=2, dead=0x01)
    np.save(os.path.join(HERE, "c3_go_dead01_256.npy"), rules)
    info = {"rules": int(len(rules)), "checked": st.checked, "qc_hits": st.qc_hits,
            "exh_calls": st.exh_calls, "seconds": round(time.time() - t0, 1)}
    print(info, flush=True)
    with open(os.path.join(HERE, "c3_counts.json"), "w") as f:
        json.dump(info, f, indent=1)


if __name__ == "__main__":
    main()
