"""The packed interpreter of the exhaustive sweep (csrc/sx_pk.h: two states
per register, 16-bit halves) against the oracle's executor (pkg/cpu/exec.go
restated in oracle/sx_oracle.c): every one of the 2,863 instructions on
random state pairs, both halves, all eight state bytes."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def pk_lib(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    out = tmp_path_factory.mktemp("pk") / "libpk.so"
    src = os.path.join(ROOT, "tests", "native", "pk_exec.cpp")
    subprocess.run([HIPCC, "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(out), src], check=True,
                   capture_output=True)
    lib = C.CDLL(str(out))
    lib.pk_exec.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64]
    return lib


def test_packed_step_matches_the_oracle(pk_lib):
    rng = np.random.default_rng(7)
    words = np.repeat(o.instruction_table(0), 64)  # 64 state pairs per instruction
    n = len(words)
    lo = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
    hi = rng.integers(0, 256, size=(n, 8), dtype=np.uint8)
    # include the edge values every flag rule turns on
    lo[::7, 0] = 0
    hi[::5, 0] = 0xFF
    lo[::11, 5] = 0x80
    want_lo = o.exec_many(lo, words)
    want_hi = o.exec_many(hi, words)
    got_lo, got_hi = lo.copy(), hi.copy()
    assert pk_lib.pk_exec(got_lo.ctypes.data, got_hi.ctypes.data, words.ctypes.data, n) == 0
    assert np.array_equal(got_lo, want_lo)
    assert np.array_equal(got_hi, want_hi)


def test_packed_step_rejects_unknown_opcodes(pk_lib):
    lo = np.zeros((1, 8), np.uint8)
    hi = np.zeros((1, 8), np.uint8)
    w = np.array([58], np.uint32)
    assert pk_lib.pk_exec(lo.ctypes.data, hi.ctypes.data, w.ctypes.data, 1) == -1
