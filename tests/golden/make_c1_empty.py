"""The empty-candidate stream of the 4,096 sampled length-2 targets (CPU).

    python tests/golden/make_c1_empty.py [threads]

C1 checks each target against the 2,863 single instructions plus the empty
sequence (11,730,944 pairs).  Go never tries the empty candidate
(pkg/search/worker.go:123), so its rules (rep_len 0: the target can be
deleted) form a stream of their own beside the first-hit rules.  For each
of the four C1 configurations this runs the oracle with the empty candidate
and complete sweeps (mode 2), checks that the other rules are exactly the
committed c1_*.npy fixtures (which were made with Go's reduced sweep: C1 has
no pair where the two sweeps disagree), and writes the empty-candidate
rules to c1_empty.npz (one RULE_DTYPE array per configuration).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as o  # noqa: E402

RUNS = [("c1_go_dead00", dict(dead=0x00)), ("c1_go_deadFF", dict(dead=0xFF)), ("c1_go_dead01", dict(dead=0x01)),
        ("c1_cuda_dead00", dict(dead=0x00, hits=1, bytes_model=1))]


def main(nthreads: int = 8):
    targets = np.load(os.path.join(HERE, "c1_targets.npy"))
    out = {}
    for name, kw in RUNS:
        rules, st = o.search(targets, exh_mode=o.EXH_COMPLETE, include_empty=True, nthreads=nthreads, **kw)
        empty = rules["rep_len"] == 0
        main_rules = np.load(os.path.join(HERE, name + ".npy"), allow_pickle=False)
        assert np.array_equal(rules[~empty], main_rules), name
        out[name] = rules[empty]
        print(name, "empty-candidate rules:", int(empty.sum()), "checked", st.checked, flush=True)
    np.savez_compressed(os.path.join(HERE, "c1_empty.npz"), **out)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
