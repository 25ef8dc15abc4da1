"""Regenerate the committed golden fixtures from the CPU oracle.

    python tests/golden/make_golden.py

Writes (all numpy .npy, no pickles):
  c1_targets.npy            4,096 length-2 targets (seed 0), shape (4096, 2)
  c1_go_dead00.npy          oracle rules, Go semantics, --dead-flags 0x00
  c1_go_deadFF.npy          ... 0xFF (exact pass + flag-relaxed pass)
  c1_go_dead01.npy          ... 0x01
  c1_cuda_dead00.npy        oracle rules, CUDA-tool semantics (all hits, 2-byte M ops)
  c1_counts.json            stage counters of each run
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


def main(only=None):
    l2 = o.enumerate_pruned(2)
    assert len(l2) == 7_633_612
    t, idx = sampling.c1_targets(l2)
    np.save(os.path.join(HERE, "c1_targets.npy"), t.astype(np.uint32))
    np.save(os.path.join(HERE, "c1_indices.npy"), idx.astype(np.int64))
    cpath = os.path.join(HERE, "c1_counts.json")
    counts = json.load(open(cpath)) if os.path.exists(cpath) else {}
    runs = [("c1_go_dead00", dict(dead=0x00)), ("c1_go_deadFF", dict(dead=0xFF)),
            ("c1_go_dead01", dict(dead=0x01)),
            ("c1_cuda_dead00", dict(dead=0x00, hits=1, bytes_model=1))]
    for name, kw in runs:
        if only and name not in only:
            continue
        t0 = time.time()
        rules, st = o.search(t, **kw)
        np.save(os.path.join(HERE, name + ".npy"), rules)
        counts[name] = {"rules": int(len(rules)), "checked": st.checked, "qc_hits": st.qc_hits,
                        "mid_pass": st.mid_pass, "exh_calls": st.exh_calls,
                        "seconds": round(time.time() - t0, 2)}
        print(name, counts[name], flush=True)
        with open(cpath, "w") as f:
            json.dump(counts, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main(sys.argv[1:] or None)
