"""6502 instruction model on the host: opcode ids, catalog, disassembly and
the assembly parser of the CLI.

Mirrors pkg/inst/instruction.go:42-122 and pkg/inst/catalog.go:105-201 (ids,
mnemonics, cycles, byte sizes, Disassemble) and the parsers of
cmd/6502opt/main.go:285-384,465-494 (parseDeadFlags, parseAssembly,
parseSingleInstruction, parseImmediate).
"""
from __future__ import annotations

from typing import List, Sequence

MNEMONICS = [
    "TAX", "TAY", "TXA", "TYA", "TSX", "TXS", "INX", "INY", "DEX", "DEY",
    "ASL A", "LSR A", "ROL A", "ROR A", "CLC", "SEC", "CLD", "SED", "CLI", "SEI", "CLV",
    "PHA", "PLA", "PHP", "PLP", "NOP",
    "LDA #n", "LDX #n", "LDY #n", "ADC #n", "SBC #n", "AND #n", "ORA #n", "EOR #n",
    "CMP #n", "CPX #n", "CPY #n",
    "LDA M", "LDX M", "LDY M", "STA M", "STX M", "STY M", "ADC M", "SBC M", "AND M",
    "ORA M", "EOR M", "CMP M", "CPX M", "CPY M", "INC M", "DEC M", "ASL M", "LSR M",
    "ROL M", "ROR M", "BIT M",
]
NAMES = [m.replace(" #n", "_IMM").replace(" M", "_M").replace(" A", "_A") for m in MNEMONICS]
OP = {n: i for i, n in enumerate(NAMES)}
OP_COUNT = 58
NUM_INSTRUCTIONS = 47 + 11 * 256

DEAD_NONE, DEAD_ALL = 0x00, 0xFF

BYTES_GO, BYTES_CUDA = 0, 1


def has_immediate(op: int) -> bool:
    return OP["LDA_IMM"] <= op <= OP["CPY_IMM"]


def word(op: int, imm: int = 0) -> int:
    """Instruction word op | imm << 16 (the layout of inst.Instruction)."""
    return (op & 0xFFFF) | ((imm & 0xFFFF) << 16)


def op_of(w: int) -> int:
    return w & 0xFFFF


def imm_of(w: int) -> int:
    return (w >> 16) & 0xFFFF


def byte_size(op: int, model: int = BYTES_GO) -> int:
    if model == BYTES_CUDA:
        return 1 if op <= OP["NOP"] else 2
    return 2 if has_immediate(op) else 1


def cycles(op: int) -> int:
    name = NAMES[op]
    if name in ("PHA", "PHP"):
        return 3
    if name in ("PLA", "PLP"):
        return 4
    if name in ("INC_M", "DEC_M", "ASL_M", "LSR_M", "ROL_M", "ROR_M"):
        return 5
    if name.endswith("_M"):
        return 3
    return 2


def seq_byte_size(seq: Sequence[int], model: int = BYTES_GO) -> int:
    return sum(byte_size(op_of(w), model) for w in seq)


def seq_cycles(seq: Sequence[int]) -> int:
    return sum(cycles(op_of(w)) for w in seq)


def disassemble(w: int) -> str:
    op = op_of(w)
    m = MNEMONICS[op]
    if not has_immediate(op):
        return m
    return m.replace("#n", "#$%02X" % (imm_of(w) & 0xFF))


def disasm_seq(seq: Sequence[int]) -> str:
    return " : ".join(disassemble(w) for w in seq)


def instruction_table() -> List[int]:
    """All 2,863 instructions in Go enumeration order
    (pkg/search/enumerator.go:14-38)."""
    out = [word(op) for op in range(OP_COUNT) if not has_immediate(op)]
    for op in range(OP_COUNT):
        if has_immediate(op):
            out.extend(word(op, imm) for imm in range(256))
    return out


def parse_dead_flags(s: str) -> int:
    low = s.lower()
    if low in ("none", ""):
        return DEAD_NONE
    if low == "all":
        return DEAD_ALL
    v = low[2:] if low.startswith("0x") else low
    try:
        n = int(v, 16)
    except ValueError:
        n = -1
    if not 0 <= n <= 0xFF:
        raise ValueError(f"invalid --dead-flags value {v!r}: use none, all, or hex (e.g. 0xFF)")
    return n


def parse_immediate(s: str) -> int:
    s = s.strip()
    if not s:
        raise ValueError("empty")
    if s.startswith("$"):
        v = int(s[1:], 16)
    elif s.startswith(("0X", "0x")):
        v = int(s[2:], 16)
    elif s.endswith("H"):
        v = int(s[:-1], 16)
    else:
        v = int(s, 10)
    if not 0 <= v <= 0xFFFF:
        raise ValueError("out of range")
    return v


def parse_single_instruction(text: str) -> int:
    text = text.strip()
    upper = text.upper()
    for op in range(OP_COUNT):
        m = MNEMONICS[op]
        if not has_immediate(op):
            if text.lower() == m.lower():
                return word(op)
            continue
        pattern = m.upper()
        n_idx = pattern.rfind("N")
        prefix, suffix = pattern[:n_idx], pattern[n_idx + 1:]
        if not upper.startswith(prefix):
            continue
        if suffix and not upper.endswith(suffix):
            continue
        val = upper[len(prefix):]
        if suffix:
            val = val[: len(val) - len(suffix)]
        try:
            v = parse_immediate(val.strip())
        except ValueError:
            continue
        return word(op, v)
    raise ValueError(f"unknown instruction: {text}")


def parse_assembly(text: str) -> List[int]:
    seq = []
    for part in text.split(":"):
        part = part.strip()
        if not part:
            continue
        try:
            seq.append(parse_single_instruction(part))
        except ValueError as e:
            raise ValueError(f"cannot parse {part!r}: {e}") from None
    if not seq:
        raise ValueError(f"no instructions parsed from {text!r}")
    return seq


def dead_flag_desc(mask: int) -> str:
    """pkg/result/output.go:77-101"""
    if mask == 0:
        return ""
    if mask == 0xFF:
        return "all"
    names = [(0x01, "C"), (0x02, "Z"), (0x04, "I"), (0x08, "D"), (0x40, "V"), (0x80, "N")]
    return ",".join(n for b, n in names if mask & b)
