"""Length-3 target ranges generated on the device (raw Go-order triple
indices, pruned by ShouldPrune) and the chunked, checkpointed enumeration of
the command line: the same rules as the explicit target list, a run
interrupted between a chunk's JSONL write and its checkpoint resumes to
exactly what an uninterrupted one writes, and --max-target 3 runs the
length-2 targets (against candidates of length 1 and 2) before the length-3
ones (search.go:26-50)."""
import json
import os
import subprocess

import numpy as np
import pytest

import oracle as o
from sx6502 import core

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "6502-optimizer_amd", "6502opt")
N = 2863
R0 = 50 * N * N + 300 * N  # targets starting LDA #$03 : LDX #..
N2 = 7_633_612              # pruned length-2 targets: --max-target 3's index space starts with them
G0 = N2 + R0                # the command line's index of raw triple R0


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


def pruned_triples(r0, n):
    tab = o.instruction_table(0)
    keep, raw = [], []
    for r in range(r0, r0 + n):
        seq = [int(tab[r // (N * N)]), int(tab[(r // N) % N]), int(tab[r % N])]
        if not o.should_prune(seq):
            keep.append(seq)
            raw.append(r)
    return np.array(keep, np.uint32), raw


def test_implicit_range_equals_explicit_targets(dev):
    n = 200000
    explicit, raw = pruned_triples(R0, n)
    print(f"{len(raw)} of {n} triples survive the pruner", flush=True)
    ri, st_i = dev.enumerate(None, target_len=3, max_cand_len=2, dead=0x01, first=R0, end=R0 + n)
    re_, st_e = dev.enumerate(explicit, target_len=3, max_cand_len=2, dead=0x01)
    assert st_i["targets"] == len(raw) == st_e["targets"]
    key = lambda r, ti: (ti, int(r["pass_"]), tuple(int(w) for w in r["rep"][: r["rep_len"]]))
    got = {key(r, int(r["target_index"])) for r in ri}
    want = {key(r, raw[int(r["target_index"])]) for r in re_}
    assert got == want and len(got) > 0
    # the rule sources are the triples their index names
    tab = o.instruction_table(0)
    for r in ri[:200]:
        t = int(r["target_index"])
        assert [int(w) for w in r["src"][:3]] == [int(tab[t // (N * N)]), int(tab[(t // N) % N]), int(tab[t % N])]


def test_all_hits_range_rules_name_their_triples(dev):
    """All-hits rules of an implicit length-3 range carry raw triple indices
    (rewritten on the device before the copy to the host)."""
    ri, st = dev.enumerate(None, target_len=3, max_cand_len=2, dead=0x01, hits=core.HITS_ALL, first=R0,
                           end=R0 + 100000)
    assert len(ri) > 100
    tab = o.instruction_table(0)
    for r in ri:
        t = int(r["target_index"])
        assert R0 <= t < R0 + 100000
        assert [int(w) for w in r["src"][:3]] == [int(tab[t // (N * N)]), int(tab[(t // N) % N]), int(tab[t % N])]


def cli(*args, env=None, ok=True):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300, env=env)
    if ok:
        assert p.returncode == 0, p.stderr + p.stdout
    return p


def test_checkpointed_enumeration_resumes_after_a_crash(tmp_path):
    chunk = 65536
    args = ["enumerate", "--max-target", "3", "--dead-flags", "0x01", "--first-target", str(G0),
            "--end-target", str(G0 + 4 * chunk), "--chunk", str(chunk)]
    full, part, ck = tmp_path / "full.jsonl", tmp_path / "part.jsonl", tmp_path / "ck.json"
    cli(*args, "--jsonl", str(full), "--checkpoint", str(tmp_path / "ck_full.json"))
    # the process ends after chunk 3's JSONL write, before its checkpoint
    p = cli(*args, "--jsonl", str(part), "--checkpoint", str(ck),
            env={**os.environ, "SX_TEST_EXIT_BEFORE_CHECKPOINT": "3"}, ok=False)
    assert p.returncode == 3, p.stderr
    c = json.loads(ck.read_text())
    assert c["next"] == G0 + 2 * chunk and c["offset"] < part.stat().st_size  # chunk 3 written, not committed
    committed = "".join(full.read_text().splitlines(True)[:c["rules"]]).encode()
    assert c["offset"] == len(committed) and part.read_bytes()[:c["offset"]] == committed
    # another configuration is refused
    p = cli(*args, "--no-mid", "--jsonl", str(part), "--checkpoint", str(ck), "--resume", ok=False)
    assert p.returncode != 0 and "another configuration" in p.stderr
    out = cli(*args, "--jsonl", str(part), "--checkpoint", str(ck), "--resume").stdout
    assert "Resuming at target" in out
    assert part.read_text() == full.read_text()
    c = json.loads(ck.read_text())
    assert c["next"] == c["end"] == G0 + 4 * chunk
    assert c["rules"] == len(full.read_text().splitlines())
    assert c["offset"] == full.stat().st_size


def test_max_target_3_runs_length_2_targets_first(dev, tmp_path):
    """Across the boundary of the two target spaces: the last 3,000 length-2
    targets (candidates of length 1 and 2), then 40,000 raw triples."""
    jl = tmp_path / "x.jsonl"
    cli("enumerate", "--max-target", "3", "--first-target", str(N2 - 3000), "--end-target", str(N2 + 40000),
        "--jsonl", str(jl), "--chunk", "7000", "--checkpoint", str(tmp_path / "ck.json"))
    lines = jl.read_text().splitlines()
    r2, _ = dev.enumerate(None, target_len=2, max_cand_len=2, first=N2 - 3000, end=N2)
    r3, _ = dev.enumerate(None, target_len=3, max_cand_len=2, first=0, end=40000)
    assert len(lines) == len(r2) + len(r3) and len(r3) > 0
    src = [json.loads(x)["source_asm"] for x in lines]
    assert all(s.count(":") == 1 for s in src[:len(r2)]) and all(s.count(":") == 2 for s in src[len(r2):])
