"""Golden fixtures for the batched length-3 configuration (C3).

    python tests/golden/make_golden_c3.py

Writes c3_targets.npy (65,536 length-3 targets drawn by rejection with
SplitMix64 seed 1, see sampling.py) and the oracle's rules for its first 512
targets against all length-1 and length-2 candidates, Go semantics
(first hit in Go order, no MidCheck), --dead-flags 0x01, complete sweeps
(oracle mode 2: Go's reduced sweep for pairs with >= 3 extra registers
replaced by a complete one) -> c3_dead01_512.npy.
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


N_SUBSET = 512


def main():
    tab = o.instruction_table(0)
    t0 = time.time()
    targets = sampling.c3_targets(tab, o.should_prune)
    path = os.path.join(HERE, "c3_targets.npy")
    if os.path.exists(path):
        assert np.array_equal(np.load(path), targets)
    np.save(path, targets)
    print("sampled", targets.shape, round(time.time() - t0, 1), flush=True)
    t0 = time.time()
    rules, st = o.search(targets[:N_SUBSET], max_cand_len=2, dead=0x01, exh_mode=o.EXH_COMPLETE, use_mid=0)
    np.save(os.path.join(HERE, f"c3_dead01_{N_SUBSET}.npy"), rules)
    info = {"rules": int(len(rules)), "checked": st.checked, "qc_hits": st.qc_hits,
            "exh_calls": st.exh_calls, "seconds": round(time.time() - t0, 1)}
    print(info, flush=True)
    with open(os.path.join(HERE, "c3_counts.json"), "w") as f:
        json.dump(info, f, indent=1)


if __name__ == "__main__":
    main()
