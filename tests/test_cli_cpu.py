"""The 6502opt command line's paths that never reach the device: verify-jsonl
lines that are skipped (cmd/6502opt/main.go:400-436), its error exits, and
argument errors.  With no parseable rule the command opens no device, so
these run anywhere the binary is built."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "6502-optimizer_amd", "6502opt")
pytestmark = pytest.mark.skipif(not os.path.exists(CLI), reason="6502opt not built")


def run(*args):
    return subprocess.run([CLI, *args], capture_output=True, text=True, timeout=60)


def test_verify_jsonl_skips_what_it_cannot_parse(tmp_path):
    f = tmp_path / "skip.jsonl"
    f.write_text("\n".join([
        "not json",                                                  # json.Unmarshal fails
        '{"source_asm":"TXA : TAX","replacement_asm":""}',           # the empty replacement
        '{"source_asm":"FOO #$01","replacement_asm":"TXA"}',         # unknown mnemonic
        "",                                                          # blank lines are not counted
        '{"replacement_asm":"TXA"}',                                 # a missing field decodes as ""
        '{"source_asm":"LDA #$1G : TAX","replacement_asm":"TXA"}',   # bad immediate
    ]) + "\n")
    p = run("verify-jsonl", str(f))
    assert p.returncode == 0, p.stderr
    err = p.stderr.splitlines()
    assert err == [
        "  [1] JSON parse error",
        '  [2] Cannot parse replacement "": no instructions parsed from ""',
        '  [3] Cannot parse source "FOO #$01": cannot parse "FOO #$01": unknown instruction: FOO #$01',
        '  [4] Cannot parse source "": no instructions parsed from ""',
        '  [5] Cannot parse source "LDA #$1G : TAX": cannot parse "LDA #$1G": unknown instruction: LDA #$1G',
    ]
    assert p.stdout.strip() == "Verification complete: 5 total, 0 passed, 0 failed, 5 skipped"


def test_verify_jsonl_errors():
    p = run("verify-jsonl", "/nonexistent/rules.jsonl")
    assert p.returncode != 0 and "no such file" in p.stderr
    p = run("verify-jsonl")
    assert p.returncode != 0 and "accepts 1 arg(s)" in p.stderr
    p = run("verify-jsonl", os.devnull, "--dead-flags", "0xZZ")
    assert p.returncode != 0 and "invalid --dead-flags" in p.stderr
    p = run("verify-jsonl", os.devnull)  # empty input: nothing to verify
    assert p.returncode == 0 and "0 total, 0 passed, 0 failed, 0 skipped" in p.stdout


def test_argument_errors():
    p = run("target", "FOO #$01")
    assert p.returncode != 0 and "failed to parse" in p.stderr
    p = run("stoke")
    assert p.returncode != 0 and "--target is required" in p.stderr
    p = run("frobnicate")
    assert p.returncode != 0 and "unknown command" in p.stderr
