"""The command line's rule files, text for text, and verify-jsonl's verdicts.

`enumerate --output` writes result.WriteJSON's format (pkg/result/output.go:
20-35: json.Encoder with two-space indent over table.Rules(), sorted by bytes
then cycles saved, pkg/result/table.go:38-52) and `--jsonl` the CUDA tool's
lines (cuda/6502_search_v2.cu:924-929).  The expected files are built here
from the oracle's rules of the same target slice with Python's json module
and the oracle's disassembler.  verify-jsonl is run on a file with passing,
failing and unparseable lines and every immediate form parseImmediate takes
($, 0x, H suffix, decimal; cmd/6502opt/main.go:465-494), with and without
dead flags; a failing rule makes the command exit non-zero (main.go:459-461).
"""
import json
import os
import subprocess

import pytest

import oracle as o

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "6502-optimizer_amd", "6502opt")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=300)


def dead_desc(mask):  # pkg/result/output.go:76-103
    if mask == 0:
        return ""
    if mask == 0xFF:
        return "all"
    return ",".join(n for b, n in ((1, "C"), (2, "Z"), (4, "I"), (8, "D"), (0x40, "V"), (0x80, "N")) if mask & b)


@pytest.mark.parametrize("dead,first,end", [(0x00, 0, 6000), (0xFF, 5_000_000, 5_004_000)])
def test_enumerate_files_match_the_formats(tmp_path, dead, first, end):
    js, jl = tmp_path / "rules.json", tmp_path / "rules.jsonl"
    p = run("enumerate", "--max-target", "2", "--dead-flags", f"0x{dead:02X}", "--first-target", str(first),
            "--end-target", str(end), "--output", str(js), "--jsonl", str(jl))
    assert p.returncode == 0, p.stderr
    targets = o.enumerate_pruned(2)[first:end]
    rules, _ = o.search(targets, dead=dead, exh_mode=o.EXH_COMPLETE, nthreads=8)
    assert len(rules) > 0
    dis = lambda ws, n: [o.disasm(int(w)) for w in ws[:n]]  # noqa: E731
    rows = [{"source": dis(r["src"], r["src_len"]), "replacement": dis(r["rep"], r["rep_len"]),
             "bytes_saved": int(r["bytes_saved"]), "cycles_saved": int(r["cycles_saved"]),
             "dead": int(r["dead_flags"])} for r in rules]
    # JSONL: one line per rule in enumeration order
    lines = jl.read_text().splitlines()
    assert len(lines) == len(rows)
    for line, r, rule in zip(lines, rows, rules):
        sb = sum(o.byte_size(int(w) & 0xFFFF) for w in rule["src"][: rule["src_len"]])
        rb = sum(o.byte_size(int(w) & 0xFFFF) for w in rule["rep"][: rule["rep_len"]])
        want = ('{"source_asm":"%s","replacement_asm":"%s","source_bytes":%d,"replacement_bytes":%d,'
                '"bytes_saved":%d,"cycles_saved":%d' % (" : ".join(r["source"]), " : ".join(r["replacement"]), sb,
                                                        rb, r["bytes_saved"], r["cycles_saved"]))
        if r["dead"]:
            want += ',"dead_flags":"0x%02X"' % r["dead"]
        assert line == want + "}"
    # JSON: table.Rules() order (stable here: ties keep enumeration order), two-space indent, trailing newline
    order = sorted(range(len(rows)), key=lambda i: (-rows[i]["bytes_saved"], -rows[i]["cycles_saved"]))
    objs = []
    for i in order:
        d = {k: rows[i][k] for k in ("source", "replacement", "bytes_saved", "cycles_saved")}
        if rows[i]["dead"]:
            d["dead_flags"] = dead_desc(rows[i]["dead"])
        objs.append(d)
    assert js.read_text() == json.dumps(objs, indent=2) + "\n"
    # and verify-jsonl passes every line under the run's mask
    v = run("verify-jsonl", str(jl), "--dead-flags", f"0x{dead:02X}")
    assert v.returncode == 0 and f"{len(lines)} total, {len(lines)} passed, 0 failed, 0 skipped" in v.stdout


def test_verify_jsonl_pass_fail_skip(tmp_path):
    f = tmp_path / "mixed.jsonl"
    f.write_text("\n".join([
        '{"source_asm":"LDA #$01 : LDA #$02","replacement_asm":"LDA #$02"}',     # 1 pass ($)
        '{"source_asm":"LDA #0x01 : LDA #0x02","replacement_asm":"LDA #2"}',     # 2 pass (0x, decimal)
        '{"source_asm":"LDX #01H : LDX #0FEH","replacement_asm":"LDX #$FE"}',    # 3 pass (H suffix)
        '{"source_asm":"CLC : LDA #$05","replacement_asm":"LDA #$05"}',          # 4 fails unless carry is dead
        '{"source_asm":"CPX M : CPY #$13","replacement_asm":"CPY #$12"}',        # 5 fails unless every flag is dead
        '{"source_asm":"TXA : TAX","replacement_asm":""}',                       # 6 skipped
        'garbage',                                                               # 7 skipped
        '{"source_asm":"txa : tax","replacement_asm":"txa"}',                    # 8 pass (mnemonics fold case)
    ]) + "\n")
    p = run("verify-jsonl", str(f))
    assert p.returncode != 0 and "2 rules failed verification" in p.stderr
    assert "Verification complete: 8 total, 4 passed, 2 failed, 2 skipped" in p.stdout
    assert "  [4] FAIL: CLC : LDA #$05 -> LDA #$05" in p.stdout
    assert "  [5] FAIL: CPX M : CPY #$13 -> CPY #$12" in p.stdout
    assert "PASS" not in p.stdout
    p = run("verify-jsonl", str(f), "--dead-flags", "0x01", "--verbose")
    assert p.returncode != 0 and "Verification complete: 8 total, 5 passed, 1 failed, 2 skipped" in p.stdout
    assert "  [4] PASS: CLC : LDA #$05 -> LDA #$05" in p.stdout and "  [3] PASS: " in p.stdout
    p = run("verify-jsonl", str(f), "--dead-flags", "all")
    assert p.returncode == 0 and "Verification complete: 8 total, 6 passed, 0 failed, 2 skipped" in p.stdout
    # the oracle's verdicts on the same pairs
    w = o.word
    clc, cpx = [w(o.CLC), w(o.LDA_IMM, 5)], [w(o.CPX_M), w(o.CPY_IMM, 0x13)]
    assert not o.exhaustive_check(clc, [w(o.LDA_IMM, 5)], 0, mode=2)
    assert o.exhaustive_check(clc, [w(o.LDA_IMM, 5)], 0x01, mode=2)
    assert not o.exhaustive_check(cpx, [w(o.CPY_IMM, 0x12)], 0x01, mode=2)
    assert o.exhaustive_check(cpx, [w(o.CPY_IMM, 0x12)], 0xFF, mode=2)
