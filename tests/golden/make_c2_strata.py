"""A stratified sample of the full length-2 enumeration with the oracle's
rule sets at --dead-flags 0x00 and 0xFF (CPU only).

    python tests/golden/make_c2_strata.py [threads]

Targets (indices into the 7,633,612 pruned length-2 targets in Go order):
every target that reads three or more extra registers on its own (the pairs
whose sweeps are largest, where Go's reduced sweep and the complete sweep can
part), plus 61,440 drawn uniformly (SplitMix64 seed 3, see sampling.py).
Rules: the oracle's worker-pool restatement (pkg/search/worker.go:119-245)
in Go semantics -- first hit in Go order, Go byte sizes, QuickCheck then
ExhaustiveCheck with no MidCheck -- with complete sweeps (oracle mode 2),
i.e. what Go's enumerate emits with its reduced sweep replaced by a complete
one.

Writes c2_strata.npz:
  index      (n,) int64    target indices, ascending
  k          (n,) uint8    extra registers each target reads on its own
  r00_*, rFF_* rule fields: t (position in `index`), rep (word), pass, dead,
             bytes, cycles
and c2_strata_counts.json (stage counters and times).
"""
import json
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))
sys.path.insert(0, HERE)

import oracle as o  # noqa: E402
import sampling  # noqa: E402

N_UNIFORM = 61_440
EXTRA_BITS = (4, 8, 16, 32, 64, 128)  # X, Y, S, M, S0, S1 read bits (verifier.go:134-143)


def own_extras(l2: np.ndarray) -> np.ndarray:
    reads = np.array([o.op_reads(op) for op in range(58)], np.int64)
    r = reads[l2[:, 0] & 0xFFFF] | reads[l2[:, 1] & 0xFFFF]
    return sum(((r & b) != 0).astype(np.uint8) for b in EXTRA_BITS).astype(np.uint8)


def strata(l2: np.ndarray) -> np.ndarray:
    k = own_extras(l2)
    idx = np.union1d(np.nonzero(k >= 3)[0], sampling.sample_indices(len(l2), N_UNIFORM, 3))
    return idx.astype(np.int64), k[idx]


def compact(rules: np.ndarray) -> dict:
    """The fields of length-2-target, single-instruction rules; t = the
    rule's target_index (the target's position in the call's list)."""
    return {"t": rules["target_index"].astype(np.uint32),
            "rep": rules["rep"][:, 0].astype(np.uint32), "pass": rules["pass_"].astype(np.uint8),
            "dead": rules["dead_flags"].astype(np.uint8), "bytes": rules["bytes_saved"].astype(np.int8),
            "cycles": rules["cycles_saved"].astype(np.int8)}


def main(nthreads: int = 8):
    l2 = o.enumerate_pruned(2)
    assert len(l2) == 7_633_612
    idx, k = strata(l2)
    targets = l2[idx]
    print(f"{len(idx)} targets, {int((k >= 3).sum())} reading >= 3 extra registers", flush=True)
    out = {"index": idx, "k": k}
    counts = {"targets": int(len(idx)), "k3": int((k == 3).sum()), "k4": int((k >= 4).sum())}
    for dead in (0x00, 0xFF):
        t0 = time.time()
        rules, st = o.search(targets, dead=dead, exh_mode=o.EXH_COMPLETE, use_mid=0, nthreads=nthreads)
        assert (rules["rep_len"] == 1).all()
        for f, v in compact(rules).items():
            out[f"r{dead:02X}_{f}"] = v
        counts[f"dead{dead:02X}"] = {"rules": int(len(rules)), "checked": st.checked, "qc_hits": st.qc_hits,
                                     "exh_calls": st.exh_calls, "seconds": round(time.time() - t0, 1)}
        print(f"dead 0x{dead:02X}:", counts[f"dead{dead:02X}"], flush=True)
    np.savez_compressed(os.path.join(HERE, "c2_strata.npz"), **out)
    with open(os.path.join(HERE, "c2_strata_counts.json"), "w") as f:
        json.dump(counts, f, indent=1)


if __name__ == "__main__":
    main(int(sys.argv[1]) if len(sys.argv) > 1 else 8)
