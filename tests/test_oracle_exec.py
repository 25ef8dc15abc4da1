"""Pins the oracle's executor (oracle/sx_oracle.c, restating pkg/cpu/exec.go).

* Known answers from pkg/cpu/exec_test.go:9-560.
* ADC/SBC over all 2^17 (A, M, C) against closed forms (exec_test.go:186-260).
* Every opcode against an independent textbook model written here from
  signed/unsigned arithmetic (not from the oracle's code), over full grids of
  the bytes it reads with every other byte random.
* The reference CUDA tool's own host executor (cuda/6502_common.h), compiled
  into oracle/_ref when the reference tree is present.
"""
import os
import subprocess

import numpy as np
import pytest

import oracle as o

A, X, Y, P, S, M, S0, S1 = range(8)
FC, FZ, FI, FD, FB, FU, FV, FN = 1, 2, 4, 8, 16, 32, 64, 128


def run(op, imm=0, **regs):
    st = o.state(**regs)
    out, _ = o.exec_op(st, op, imm)
    return dict(zip("a x y p s m s0 s1".split(), out))


def test_opcode_and_instruction_counts():
    assert len(o.OPS) == 58                              # search_test.go:201-205
    tab = o.instruction_table(0)
    nonimm = [w for w in tab if not (26 <= (w & 0xFFFF) <= 36)]
    assert len(nonimm) == 47 and len(tab) == 47 + 11 * 256   # search_test.go:9-22
    assert sorted(tab.tolist()) == sorted(o.instruction_table(1).tolist())


def test_transfers():  # exec_test.go:9-43
    r = run(o.TAX, a=0x42)
    assert r["x"] == 0x42 and not r["p"] & FZ
    r = run(o.TAY, a=0)
    assert r["y"] == 0 and r["p"] & FZ
    r = run(o.TXA, x=0x80)
    assert r["a"] == 0x80 and r["p"] & FN
    assert run(o.TYA, y=0x55)["a"] == 0x55


def test_txs_tsx():  # exec_test.go:45-64
    r = run(o.TXS, x=0xFD)
    assert r["s"] == 0xFD and r["p"] & (FN | FZ) == 0
    r = run(o.TSX, s=0x80)
    assert r["x"] == 0x80 and r["p"] & FN


def test_inc_dec():  # exec_test.go:66-102
    r = run(o.INX, x=0xFF)
    assert r["x"] == 0 and r["p"] & FZ
    r = run(o.INY, y=0x7F)
    assert r["y"] == 0x80 and r["p"] & FN
    r = run(o.DEX, x=0)
    assert r["x"] == 0xFF and r["p"] & FN
    r = run(o.DEY, y=1)
    assert r["y"] == 0 and r["p"] & FZ


@pytest.mark.parametrize("a,imm,c,wa,wc,wv,wz,wn", [  # exec_test.go:115-121
    (0x10, 0x20, 0, 0x30, 0, 0, 0, 0), (0x10, 0x20, 1, 0x31, 0, 0, 0, 0),
    (0xFF, 0x01, 0, 0x00, 1, 0, 1, 0), (0x7F, 0x01, 0, 0x80, 0, 1, 0, 1),
    (0x80, 0x80, 0, 0x00, 1, 1, 1, 0), (0x40, 0x20, 0, 0x60, 0, 0, 0, 0),
    (0x00, 0x00, 0, 0x00, 0, 0, 1, 0)])
def test_adc_table(a, imm, c, wa, wc, wv, wz, wn):
    r = run(o.ADC_IMM, imm, a=a, p=c)
    assert (r["a"], bool(r["p"] & FC), bool(r["p"] & FV), bool(r["p"] & FZ),
            bool(r["p"] & FN)) == (wa, bool(wc), bool(wv), bool(wz), bool(wn))


@pytest.mark.parametrize("a,imm,c,wa,wc,wv", [  # exec_test.go:159-163
    (0x30, 0x10, 1, 0x20, 1, 0), (0x30, 0x10, 0, 0x1F, 1, 0), (0x00, 0x01, 1, 0xFF, 0, 0),
    (0x80, 0x01, 1, 0x7F, 1, 1), (0x7F, 0xFF, 1, 0x80, 0, 1)])
def test_sbc_table(a, imm, c, wa, wc, wv):
    r = run(o.SBC_IMM, imm, a=a, p=c)
    assert (r["a"], bool(r["p"] & FC), bool(r["p"] & FV)) == (wa, bool(wc), bool(wv))


def _grid17():
    a, m, c = np.meshgrid(np.arange(256), np.arange(256), np.arange(2), indexing="ij")
    return a.ravel().astype(np.int64), m.ravel().astype(np.int64), c.ravel().astype(np.int64)


def test_adc_sbc_exhaustive_closed_form():  # exec_test.go:186-260
    a, m, c = _grid17()
    st = np.zeros((len(a), 8), np.uint8)
    st[:, A], st[:, P] = a, c
    out = o.exec_many(st, (o.ADC_IMM | (m << 16)).astype(np.uint32))
    r16 = a + m + c
    wa = r16 & 0xFF
    assert np.array_equal(out[:, A], wa)
    assert np.array_equal(out[:, P] & FC != 0, r16 > 0xFF)
    assert np.array_equal(out[:, P] & FZ != 0, wa == 0)
    assert np.array_equal(out[:, P] & FN != 0, wa & 0x80 != 0)
    assert np.array_equal(out[:, P] & FV != 0, ((a ^ wa) & (m ^ wa) & 0x80) != 0)
    out = o.exec_many(st, (o.SBC_IMM | (m << 16)).astype(np.uint32))
    r16 = (a - m - (1 - c)) & 0xFFFF
    wa = r16 & 0xFF
    assert np.array_equal(out[:, A], wa)
    assert np.array_equal(out[:, P] & FC != 0, r16 <= 0xFF)
    assert np.array_equal(out[:, P] & FZ != 0, wa == 0)
    assert np.array_equal(out[:, P] & FN != 0, wa & 0x80 != 0)
    assert np.array_equal(out[:, P] & FV != 0, ((a ^ wa) & (a ^ m) & 0x80) != 0)


def test_compare_shift_logic_stack_bit_memory_flags():  # exec_test.go:262-560
    r = run(o.CMP_IMM, 0x30, a=0x30)
    assert r["p"] & FZ and r["p"] & FC and r["a"] == 0x30
    r = run(o.CMP_IMM, 0x40, a=0x30)
    assert not r["p"] & FZ and not r["p"] & FC and r["p"] & FN
    r = run(o.CMP_IMM, 0x10, a=0x30)
    assert not r["p"] & FZ and r["p"] & FC
    assert run(o.CPX_IMM, 0x42, x=0x42)["p"] & FZ
    assert not run(o.CPY_IMM, 0x01, y=0)["p"] & FC
    r = run(o.ASL_A, a=0x81)
    assert r["a"] == 0x02 and r["p"] & FC
    r = run(o.LSR_A, a=0x03)
    assert r["a"] == 0x01 and r["p"] & FC
    r = run(o.ROL_A, a=0x80, p=FC)
    assert r["a"] == 0x01 and r["p"] & FC
    r = run(o.ROR_A, a=0x01, p=FC)
    assert r["a"] == 0x80 and r["p"] & FC and r["p"] & FN
    assert run(o.AND_IMM, 0x0F, a=0xFF)["a"] == 0x0F
    r = run(o.ORA_IMM, 0xF0, a=0x0F)
    assert r["a"] == 0xFF and r["p"] & FN
    r = run(o.EOR_IMM, 0xFF, a=0xFF)
    assert r["a"] == 0 and r["p"] & FZ
    st, _ = o.exec_op(o.state(a=0x42, s=0xFD), o.PHA)
    assert st[S] == 0xFC and st[S0] == 0x42
    st[A] = 0
    st, _ = o.exec_op(st, o.PLA)
    assert st[A] == 0x42 and st[S] == 0xFD
    st, _ = o.exec_op(o.state(p=FC | FN, s=0xFD), o.PHP)
    assert st[S0] == FC | FN | FB | FU
    st[P] = 0
    st, _ = o.exec_op(st, o.PLP)
    assert st[P] == FC | FN | FB | FU
    r = run(o.BIT_M, a=0, m=0xC0)
    assert r["p"] & FN and r["p"] & FV and r["p"] & FZ
    r = run(o.BIT_M, a=0x0F, m=0x0F)
    assert not r["p"] & (FN | FV | FZ)
    assert run(o.LDA_M, m=0x42)["a"] == 0x42
    assert run(o.STA_M, a=0x55)["m"] == 0x55
    r = run(o.INC_M, m=0xFF)
    assert r["m"] == 0 and r["p"] & FZ
    r = run(o.DEC_M, m=0)
    assert r["m"] == 0xFF and r["p"] & FN
    r = run(o.ASL_M, m=0x81)
    assert r["m"] == 0x02 and r["p"] & FC
    st = o.state(p=0xFF)
    for op, bit, on in ((o.CLC, FC, 0), (o.SEC, FC, 1), (o.CLV, FV, 0), (o.CLD, FD, 0), (o.SED, FD, 1)):
        st, _ = o.exec_op(st, op)
        assert bool(st[P] & bit) == bool(on)
    # two-deep LIFO (exec_test.go:519-540)
    st = o.state(a=0x11, s=0xFF)
    st, _ = o.exec_op(st, o.PHA)
    st[A] = 0x22
    st, _ = o.exec_op(st, o.PHA)
    assert st[S] == 0xFD
    st, _ = o.exec_op(st, o.PLA)
    assert st[A] == 0x22
    st, _ = o.exec_op(st, o.PLA)
    assert st[A] == 0x11 and st[S] == 0xFF
    # immediate loads (exec_test.go:542-560)
    st = o.state()
    st, _ = o.exec_op(st, o.LDA_IMM, 0)
    assert st[A] == 0 and st[P] & FZ
    st, _ = o.exec_op(st, o.LDA_IMM, 0x80)
    assert st[A] == 0x80 and st[P] & FN
    st, _ = o.exec_op(st, o.LDX_IMM, 0xFF)
    assert st[X] == 0xFF and st[P] & FN
    st, _ = o.exec_op(st, o.LDY_IMM, 0x01)
    assert st[Y] == 1 and not st[P] & (FN | FZ)


from textbook import model  # noqa: E402  (independent textbook model of every opcode)


def _inputs_for(op, rng):
    """Full grid over the two bytes the op reads most (and carry), every other
    byte random."""
    main = {"X": X, "Y": Y}
    name = o.OPS[op]
    r1 = A
    for k, v in main.items():
        if k in name[:3] and name[:3] not in ("TAX", "TAY", "STA"):
            r1 = v
    if name in ("TSX",):
        r1 = S
    if name in ("PLA", "PLP"):
        r1 = S0
    g1, g2, c = _grid17()
    st = rng.integers(0, 256, size=(len(g1), 8), dtype=np.int64)
    st[:, r1] = g1
    st[:, P] = (st[:, P] & ~1) | c
    if name.endswith("_IMM"):
        imm = g2
    else:
        imm = np.zeros_like(g2)
        st[:, M if r1 != M else A] = g2
    return st.astype(np.uint8), imm


@pytest.mark.parametrize("op", range(58), ids=o.OPS)
def test_every_opcode_against_textbook_model(op):
    rng = np.random.default_rng(op + 1)
    st, imm = _inputs_for(op, rng)
    got = o.exec_many(st, (op | (imm << 16)).astype(np.uint32))
    want = model(op, st, imm)
    bad = np.nonzero((got != want).any(axis=1))[0]
    assert len(bad) == 0, (o.OPS[op], st[bad[0]], imm[bad[0]], got[bad[0]], want[bad[0]])


def test_cycles_and_bytes():
    # catalog.go:105-201 and cuda/6502_search_v2.cu:437-449
    for op in range(58):
        name = o.OPS[op]
        want_go = 2 if name.endswith("_IMM") else 1
        want_cuda = 1 if op <= o.NOP else 2
        assert o.byte_size(op, 0) == want_go and o.byte_size(op, 1) == want_cuda
        if name in ("PHA", "PHP"):
            cyc = 3
        elif name in ("PLA", "PLP"):
            cyc = 4
        elif name in ("INC_M", "DEC_M", "ASL_M", "LSR_M", "ROL_M", "ROR_M"):
            cyc = 5
        elif name.endswith("_M"):
            cyc = 3
        else:
            cyc = 2
        assert o.cycles(op) == cyc


def test_disassembly():  # catalog.go:59-85
    assert o.disasm(o.word(o.LDA_IMM, 0x42)) == "LDA #$42"
    assert o.disasm(o.word(o.CPY_IMM, 0x0A)) == "CPY #$0A"
    assert o.disasm(o.word(o.ASL_A)) == "ASL A"
    assert o.disasm(o.word(o.BIT_M)) == "BIT M"


REF_EXEC = os.path.join(os.path.dirname(o.__file__), "_ref", "ref_exec")


@pytest.mark.skipif(not os.path.exists(REF_EXEC), reason="oracle/_ref not built (reference tree absent)")
def test_against_reference_header_executor():
    """Every one of the 2,863 instructions on the 40 QC/Mid vectors and on
    2,000 random states, through the reference CUDA tool's host executor."""
    rng = np.random.default_rng(7)
    vecs = np.concatenate([o.qc_vectors(), o.mid_vectors(),
                           rng.integers(0, 256, (2000, 8)).astype(np.uint8)])
    tab = o.instruction_table(1)
    st = np.repeat(vecs[None], len(tab), axis=0).reshape(-1, 8)
    words = np.repeat(tab, len(vecs))
    rec = np.zeros((len(st), 12), np.uint8)
    rec[:, :8] = st
    rec[:, 8:10] = (words & 0xFFFF).astype("<u2").view(np.uint8).reshape(-1, 2)
    rec[:, 10:12] = (words >> 16).astype("<u2").view(np.uint8).reshape(-1, 2)
    out = subprocess.run([REF_EXEC], input=rec.tobytes(), capture_output=True, check=True).stdout
    ref = np.frombuffer(out, np.uint8).reshape(-1, 8)
    got = o.exec_many(st, words)
    assert np.array_equal(got, ref)
