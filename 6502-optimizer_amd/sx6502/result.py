"""Rule records and the two output formats.

* ``write_json``: the Go CLI's indented JSON array (pkg/result/output.go:12-35)
  after ``Table.Rules`` ordering, bytes saved then cycles saved, descending
  (pkg/result/table.go:38-51; Go's sort is unstable, ties here keep target
  order).
* ``write_jsonl``: the CUDA tool's one-object-per-line schema
  (cuda/6502_search_v2.cu:912-930).
"""
from __future__ import annotations

import io
import json
from dataclasses import dataclass
from typing import Iterable, List

import numpy as np

from . import isa


@dataclass(frozen=True)
class Rule:
    source: tuple
    replacement: tuple
    bytes_saved: int
    cycles_saved: int
    dead_flags: int = 0
    pass_: int = 1
    target_index: int = 0


def rules_from_array(arr: np.ndarray) -> List[Rule]:
    out = []
    for r in arr:
        out.append(Rule(tuple(int(w) for w in r["src"][: r["src_len"]]),
                        tuple(int(w) for w in r["rep"][: r["rep_len"]]),
                        int(r["bytes_saved"]), int(r["cycles_saved"]), int(r["dead_flags"]),
                        int(r["pass_"]), int(r["target_index"])))
    return out


def sorted_rules(rules: Iterable[Rule]) -> List[Rule]:
    return sorted(rules, key=lambda r: (-r.bytes_saved, -r.cycles_saved))


def write_json(f, rules: Iterable[Rule]) -> None:
    j = []
    for r in sorted_rules(rules):
        d = {"source": [isa.disassemble(w) for w in r.source],
             "replacement": [isa.disassemble(w) for w in r.replacement],
             "bytes_saved": r.bytes_saved, "cycles_saved": r.cycles_saved}
        desc = isa.dead_flag_desc(r.dead_flags)
        if desc:
            d["dead_flags"] = desc
        j.append(d)
    f.write(json.dumps(j, indent=2, separators=(",", ": ")))
    f.write("\n")


def jsonl_line(r: Rule, bytes_model: int = isa.BYTES_CUDA) -> str:
    sb = isa.seq_byte_size(r.source, bytes_model)
    rb = isa.seq_byte_size(r.replacement, bytes_model)
    s = ('{"source_asm":"%s","replacement_asm":"%s","source_bytes":%d,"replacement_bytes":%d,'
         '"bytes_saved":%d,"cycles_saved":%d' % (isa.disasm_seq(r.source), isa.disasm_seq(r.replacement),
                                                   sb, rb, r.bytes_saved, r.cycles_saved))
    if r.dead_flags:
        s += ',"dead_flags":"0x%02X"' % r.dead_flags
    return s + "}"


def write_jsonl(f, rules: Iterable[Rule], bytes_model: int = isa.BYTES_CUDA) -> None:
    buf = io.StringIO()
    for r in rules:
        buf.write(jsonl_line(r, bytes_model))
        buf.write("\n")
    f.write(buf.getvalue())


def read_json(f) -> list:
    return json.load(f)


def canonical_set(arr: np.ndarray) -> set:
    """Order-free identity of a rule set: (source, replacement, dead flags,
    pass, bytes, cycles) tuples."""
    out = set()
    for r in arr:
        out.add((tuple(int(w) for w in r["src"][: r["src_len"]]),
                 tuple(int(w) for w in r["rep"][: r["rep_len"]]),
                 int(r["dead_flags"]), int(r["pass_"]), int(r["bytes_saved"]), int(r["cycles_saved"])))
    return out


def canonical_digest(arr: np.ndarray) -> str:
    """SHA-256 over the canonical rule records sorted by (target, pass) and
    replacement, independent of the device's emission order."""
    import hashlib
    keys = np.zeros(len(arr), dtype=[("src", "<u4", (4,)), ("rep", "<u4", (4,)), ("src_len", "u1"),
                                     ("rep_len", "u1"), ("dead_flags", "u1"), ("pass_", "u1"),
                                     ("bytes_saved", "<i2"), ("cycles_saved", "<i2")])
    for name in keys.dtype.names:
        keys[name] = arr[name]
    order = np.lexsort(tuple(keys["rep"][:, i] for i in range(3, -1, -1)) +
                       (keys["pass_"],) + tuple(keys["src"][:, i] for i in range(3, -1, -1)))
    return hashlib.sha256(keys[order].tobytes()).hexdigest()
