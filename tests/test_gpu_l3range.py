"""Length-3 target ranges generated on the device (raw Go-order triple
indices, pruned by ShouldPrune) and the chunked, checkpointed enumeration of
the command line: the same rules as the explicit target list, and a resumed
run writes what an uninterrupted one does."""
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


def cli(*args):
    p = subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0, p.stderr + p.stdout
    return p


def test_checkpointed_enumeration_resumes(tmp_path):
    chunk = 65536
    full, part, ck = tmp_path / "full.jsonl", tmp_path / "part.jsonl", tmp_path / "ck.json"
    cli("enumerate", "--max-target", "3", "--dead-flags", "0x01", "--first-target", str(R0),
        "--end-target", str(R0 + 3 * chunk), "--jsonl", str(full), "--checkpoint", str(tmp_path / "ck_full.json"),
        "--chunk", str(chunk))
    # an interrupted run: one chunk done, the checkpoint names the whole range
    cli("enumerate", "--max-target", "3", "--dead-flags", "0x01", "--first-target", str(R0),
        "--end-target", str(R0 + chunk), "--jsonl", str(part), "--checkpoint", str(ck), "--chunk", str(chunk))
    c = json.loads(ck.read_text())
    assert c["next"] == R0 + chunk
    c["end"] = R0 + 3 * chunk
    ck.write_text(json.dumps(c))
    out = cli("enumerate", "--max-target", "3", "--dead-flags", "0x01", "--jsonl", str(part), "--checkpoint",
              str(ck), "--chunk", str(chunk), "--resume").stdout
    assert "Resuming at target" in out
    assert part.read_text() == full.read_text()
    c = json.loads(ck.read_text())
    assert c["next"] == c["end"] == R0 + 3 * chunk
    assert c["rules"] == len(full.read_text().splitlines())
