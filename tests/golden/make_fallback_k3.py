"""The CUDA tool's CPU-fallback pairs of the published full length-2 run,
with their verdicts under the reduced and the complete sweep (CPU only).

    python tests/golden/make_fallback_k3.py

The published run (README.md:127-137, cuda/6502_search_v2.cu:848-977) sends
every QuickCheck + MidCheck survivor that reads three or more extra
registers to a CPU sweep over 32 representative values per register
(cu:650,671-736,945-974; Go's exhaustiveReducedSweep, verifier.go:377-415).
This script enumerates those pairs with the oracle alone (4,902 of them,
README.md:135), decides each under the oracle's reduced sweep (mode 0) and
its complete sweep by projection (mode 2), records a disagreeing domain
state for every pair a sweep rejects, and brute-forces the whole 2^33-state
domain (mode 1) for a fixed sample of pairs.  Nothing here calls the
product library.

Writes c2_cuda_fallback_k3.npz:
  pairs     (4902, 3) uint32   target words t0, t1 and the candidate word
  reduced   (4902,)   bool     passes the reduced sweep
  complete  (4902,)   bool     passes the complete sweep
  cex       (4902, 8) uint8    complete sweep's counterexample (zeros where it passes)
  rcex      (4902, 8) uint8    reduced sweep's counterexample (zeros where it passes)
  brute_idx (n,)      int64    pairs brute-forced over the whole domain
  brute     (n,)      bool     their brute-force verdicts
"""
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

import oracle as o  # noqa: E402

OUT = os.path.join(HERE, "c2_cuda_fallback_k3.npz")
N_BRUTE_PASS, N_BRUTE_FAIL = 40, 8


def main(nthreads: int = 8):
    t0 = time.time()
    pairs = o.cuda_fallback_pairs(3, nthreads)
    print(f"{len(pairs)} fallback pairs ({time.time() - t0:.1f} s)", flush=True)

    def verdict(i):
        t, c = pairs[i, :2].tolist(), pairs[i, 2:].tolist()
        red, rcex = o.exhaustive_cex(t, c, 0, o.EXH_GO)
        com, ccex = o.exhaustive_cex(t, c, 0, o.EXH_COMPLETE)
        z = np.zeros(8, np.uint8)
        return red, com, z if ccex is None else ccex, z if rcex is None else rcex

    with ThreadPoolExecutor(nthreads) as ex:
        res = list(ex.map(verdict, range(len(pairs))))
    reduced = np.array([r[0] for r in res])
    complete = np.array([r[1] for r in res])
    print(f"reduced {reduced.sum()} complete {complete.sum()} ({time.time() - t0:.1f} s)", flush=True)
    # brute force: evenly spaced complete passes and reduced-only passes
    rng = np.random.default_rng(0)
    ip = np.nonzero(complete)[0]
    iq = np.nonzero(reduced & ~complete)[0]
    idx = np.concatenate([np.sort(rng.choice(ip, N_BRUTE_PASS, replace=False)),
                          np.sort(rng.choice(iq, N_BRUTE_FAIL, replace=False))])
    brute = []
    for n, i in enumerate(idx):
        ok, _ = o.exhaustive_brute(pairs[i, :2].tolist(), pairs[i, 2:].tolist(), 0, nthreads)
        brute.append(ok)
        print(f"  brute {n + 1}/{len(idx)}: pair {i} -> {ok} ({time.time() - t0:.0f} s)", flush=True)
    np.savez_compressed(OUT, pairs=pairs, reduced=reduced, complete=complete,
                        cex=np.array([r[2] for r in res], np.uint8), rcex=np.array([r[3] for r in res], np.uint8),
                        brute_idx=idx.astype(np.int64), brute=np.array(brute))
    print("wrote", OUT)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
