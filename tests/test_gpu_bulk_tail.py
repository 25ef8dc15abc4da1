"""The first-hit tail gives the same rules whenever it starts.

Go takes, per target and pass, the first candidate in Go order that passes
the exhaustive check (pkg/search/worker.go:96-107).  The library reaches it
by rounds (each sweeps the lowest-ranked untested Mid survivor of every open
(target, pass)) and then one batch that sweeps every untested survivor and
keeps the lowest passing rank.  SX_BULK_AFTER picks when the batch starts:
unset, at once while the survivors number at most the (target, pass)
entries and after one round otherwise; N, after N rounds; -1, once the
untested survivors are few against the open entries.  Rules must be the same
byte for byte.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
MODES = (None, "0", "1", "-1")


def _open(mode):
    old = os.environ.get("SX_BULK_AFTER")
    if mode is None:
        os.environ.pop("SX_BULK_AFTER", None)
    else:
        os.environ["SX_BULK_AFTER"] = mode
    try:
        return core.Device(0)
    finally:
        if old is None:
            os.environ.pop("SX_BULK_AFTER", None)
        else:
            os.environ["SX_BULK_AFTER"] = old


@pytest.fixture(scope="module")
def devs():
    d = {m: _open(m) for m in MODES}
    yield d
    for x in d.values():
        x.close()


@pytest.mark.parametrize("dead", [0x00, 0x01])
def test_full_length2_run(devs, dead):
    print(f"bulk tail, full length-2 run, dead 0x{dead:02X}", flush=True)
    ref, st = devs[None].enumerate(None, dead=dead)
    for m in MODES[1:]:
        r, s = devs[m].enumerate(None, dead=dead)
        assert s["rules"] == st["rules"] and r.tobytes() == ref.tobytes(), m
    # unset: the batch at once (5.4M survivors for 15.3M (target, pass) entries)
    assert st["exh_rounds"] == 1


def test_length3_sample(devs):
    print("bulk tail, length-3 sample", flush=True)
    t = np.load(os.path.join(GOLD, "c3_targets.npy"))[:8192]
    ref, st = devs[None].enumerate(t, target_len=3, max_cand_len=2, dead=0x01)
    for m in MODES[1:]:
        r, s = devs[m].enumerate(t, target_len=3, max_cand_len=2, dead=0x01)
        assert s["rules"] == st["rules"] and r.tobytes() == ref.tobytes(), m
