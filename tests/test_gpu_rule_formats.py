"""Rule records reach the host the same whichever way the targets came.

A first-hit enumeration of length-2 targets against single instructions
sends each rule to the host as 3 words, 12 bytes (k_pack_rules), in pieces whose copies
sx_fetch_rules unpacks as they land; the targets are the handle's own list
(copied on the device) or the same list passed by the caller (uploaded).
The records must be identical byte for byte, at both byte models and both
dead-flag masks, and fetches of sub-ranges must return the same records.
"""
import os

import numpy as np
import pytest

from sx6502 import core

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = core.Device(0)
    yield d
    d.close()


@pytest.fixture(scope="module")
def targets(dev):
    return dev.pruned_targets()


@pytest.mark.parametrize("dead,model", [(0x00, 0), (0xFF, 0), (0x00, 1)])
def test_device_list_equals_caller_list(dev, targets, dead, model):
    print(f"rule forms, dead 0x{dead:02X}, byte model {model}", flush=True)
    a, sa = dev.enumerate(None, dead=dead, bytes_model=model)
    b, sb = dev.enumerate(targets, dead=dead, bytes_model=model)
    assert sa["rules"] == sb["rules"] == len(a) == len(b)
    assert a.tobytes() == b.tobytes()
    # the fields the 3-word form leaves out
    assert np.all(a["src_len"] == 2) and np.all(a["rep_len"] == 1)
    assert np.all(a["src"][:, 2:] == 0) and np.all(a["rep"][:, 1:] == 0)


def test_fetch_of_sub_ranges(dev):
    print("rule forms, fetch of sub-ranges", flush=True)
    a, st = dev.enumerate(None, dead=0x00)
    n = len(a)
    for off, cnt in ((0, 1), (65_535, 3), (n // 2, 100_000), (n - 7, 7), (n - 1, 10)):
        part = dev.fetch_rules(off, cnt)
        assert part.tobytes() == a[off: off + cnt].tobytes(), (off, cnt)


def test_segments(dev):
    # 2^20-target segments: eight segments, each one's rule copies queued
    # while the next one runs, the last one's before the final wait
    print("rule forms, 2^20-target segments", flush=True)
    old = os.environ.get("SX_SEG_LOG")
    os.environ["SX_SEG_LOG"] = "20"
    try:
        d = core.Device(0)
    finally:
        if old is None:
            del os.environ["SX_SEG_LOG"]
        else:
            os.environ["SX_SEG_LOG"] = old
    try:
        for dead in (0x00, 0xFF):
            a, sa = dev.enumerate(None, dead=dead)
            b, sb = d.enumerate(None, dead=dead)
            assert sa["rules"] == sb["rules"] and a.tobytes() == b.tobytes(), dead
        n = len(b)
        for off, cnt in ((0, 70_000), (n // 3, 200_000), (n - 5, 5)):
            assert d.fetch_rules(off, cnt).tobytes() == b[off: off + cnt].tobytes(), (off, cnt)
    finally:
        d.close()
