"""Independent textbook model of the 58 opcodes, for tests.

Each opcode is defined from the 6502 data sheet's arithmetic (signed
overflow as a range test, borrow as a negative difference), vectorised over
numpy arrays of states (A X Y P S M S0 S1), without reference to the oracle's
code.  The executor tests pin the oracle against it; the reduced-sweep tests
run whole sweeps with it.
"""
import numpy as np

import oracle as o

A, X, Y, P, S, M, S0, S1 = range(8)
FC, FZ, FI, FD, FB, FU, FV, FN = 1, 2, 4, 8, 16, 32, 64, 128


def _s8(v):
    return np.where(v >= 128, v - 256, v)


def _nz(p, v):
    p = p & ~(FN | FZ)
    return p | np.where(v == 0, FZ, 0) | np.where(v >= 128, FN, 0)


def model(op, st, imm):
    st = st.astype(np.int64).copy()
    a, x, y, p, s, m, s0, s1 = (st[:, i].copy() for i in range(8))
    name = o.OPS[op]
    src = {"IMM": imm}

    def operand():
        return imm if name.endswith("_IMM") else m

    def setc(p, cond):
        return (p & ~FC) | np.where(cond, FC, 0)

    if name in ("TAX", "TAY", "TXA", "TYA", "TSX"):
        dst, srcreg = {"TAX": (X, A), "TAY": (Y, A), "TXA": (A, X), "TYA": (A, Y), "TSX": (X, S)}[name]
        st[:, dst] = st[:, srcreg]
        st[:, P] = _nz(p, st[:, dst])
    elif name == "TXS":
        st[:, S] = x
    elif name in ("INX", "INY", "DEX", "DEY"):
        r = X if name[2] == "X" else Y
        st[:, r] = (st[:, r] + (1 if name[0] == "I" else -1)) % 256
        st[:, P] = _nz(p, st[:, r])
    elif name in ("ASL_A", "LSR_A", "ROL_A", "ROR_A", "ASL_M", "LSR_M", "ROL_M", "ROR_M"):
        r = A if name.endswith("_A") else M
        v = st[:, r]
        cin = p & 1
        if name.startswith("ASL"):
            res, cout = (v * 2) % 256, v >= 128
        elif name.startswith("LSR"):
            res, cout = v // 2, v % 2 == 1
        elif name.startswith("ROL"):
            res, cout = (v * 2 + cin) % 256, v >= 128
        else:
            res, cout = v // 2 + 128 * cin, v % 2 == 1
        st[:, r] = res
        st[:, P] = _nz(setc(p, cout), res)
    elif name in ("CLC", "SEC", "CLD", "SED", "CLI", "SEI", "CLV"):
        bit = {"C": FC, "D": FD, "I": FI, "V": FV}[name[2]]
        st[:, P] = (p | bit) if name[0] == "S" else (p & ~bit)
    elif name == "PHA" or name == "PHP":
        val = a if name == "PHA" else (p | 0x30)
        st[:, S1], st[:, S0], st[:, S] = s0, val, (s - 1) % 256
    elif name == "PLA" or name == "PLP":
        val = s0
        st[:, S0], st[:, S1], st[:, S] = s1, 0, (s + 1) % 256
        if name == "PLA":
            st[:, A] = val
            st[:, P] = _nz(p, val)
        else:
            st[:, P] = val
    elif name == "NOP":
        pass
    elif name[:3] in ("LDA", "LDX", "LDY"):
        r = {"A": A, "X": X, "Y": Y}[name[2]]
        st[:, r] = operand()
        st[:, P] = _nz(p, st[:, r])
    elif name[:3] in ("STA", "STX", "STY"):
        st[:, M] = st[:, {"A": A, "X": X, "Y": Y}[name[2]]]
    elif name[:3] == "ADC":
        v = operand()
        total = a + v + (p & 1)
        signed = _s8(a) + _s8(v) + (p & 1)
        res = total % 256
        p2 = setc(p, total > 255) & ~FV
        p2 = p2 | np.where((signed < -128) | (signed > 127), FV, 0)
        st[:, A] = res
        st[:, P] = _nz(p2, res)
    elif name[:3] == "SBC":
        v = operand()
        borrow = 1 - (p & 1)
        diff = a - v - borrow
        signed = _s8(a) - _s8(v) - borrow
        res = diff % 256
        p2 = setc(p, diff >= 0) & ~FV
        p2 = p2 | np.where((signed < -128) | (signed > 127), FV, 0)
        st[:, A] = res
        st[:, P] = _nz(p2, res)
    elif name[:3] in ("AND", "ORA", "EOR"):
        v = operand()
        res = {"AND": a & v, "ORA": a | v, "EOR": a ^ v}[name[:3]]
        st[:, A] = res
        st[:, P] = _nz(p, res)
    elif name[:3] in ("CMP", "CPX", "CPY"):
        reg = {"CMP": a, "CPX": x, "CPY": y}[name[:3]]
        v = operand()
        st[:, P] = _nz(setc(p, reg >= v), (reg - v) % 256)
    elif name in ("INC_M", "DEC_M"):
        st[:, M] = (m + (1 if name == "INC_M" else -1)) % 256
        st[:, P] = _nz(p, st[:, M])
    elif name == "BIT_M":
        p2 = p & ~(FN | FV | FZ)
        st[:, P] = p2 | (m & 0xC0) | np.where((a & m) == 0, FZ, 0)
    else:
        raise AssertionError(name)
    return st.astype(np.uint8)
