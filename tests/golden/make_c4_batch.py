"""65,536 pruned length-4 targets for the batched length-4 workload (CPU).

    python tests/golden/make_c4_batch.py

Drawn uniformly (numpy PCG64, seed 5) from the 2,863 instructions, kept when
the oracle's pruner (pkg/search/pruner.go) keeps them: the length-4 analogue
of the C3 sample.  Writes c4_batch_targets.npy (uint32, 65,536 x 4).
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as o  # noqa: E402


def main(n: int = 65536, seed: int = 5):
    rng = np.random.default_rng(seed)
    tab = o.instruction_table(0)
    out = []
    while len(out) < n:
        t = [int(tab[i]) for i in rng.integers(0, len(tab), 4)]
        if not o.should_prune(t):
            out.append(t)
    np.save(os.path.join(HERE, "c4_batch_targets.npy"), np.array(out, np.uint32))


if __name__ == "__main__":
    main()
