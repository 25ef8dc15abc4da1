"""The package's pkg/search and pkg/result mirror (sx6502.search,
sx6502.result) on the device.

search.run over the 4,096 sampled length-2 targets returns the committed C1
rule sets (Go semantics at 0x00 and 0xFF); the predicates agree with the
oracle; search_single answers the README example; result.write_json and
write_jsonl write the same text as the command line for the same target
slice.
"""
import io
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle as o
from sx6502 import isa, result, search

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = os.path.join(ROOT, "tests", "golden")
CLI = os.path.join(ROOT, "6502-optimizer_amd", "6502opt")


def gold(name):
    return np.load(os.path.join(GOLD, name + ".npy"), allow_pickle=False)


@pytest.mark.parametrize("dead,fixture", [(0x00, "c1_go_dead00"), (0xFF, "c1_go_deadFF")])
def test_run_reproduces_the_c1_rules(dead, fixture):
    cfg = search.Config(max_target_len=2, dead_flags=dead, targets=gold("c1_targets"))
    rules = search.run(cfg)
    want = result.rules_from_array(gold(fixture))
    key = lambda r: (r.target_index, r.pass_)  # noqa: E731
    assert sorted(rules, key=key) == sorted(want, key=key)
    assert cfg.stats[2]["targets"] == 4096


def test_predicates_match_the_oracle():
    W = o.word
    pairs = [([W(o.TXA), W(o.TAX)], [W(o.TXA)], 0), ([W(o.CLC), W(o.LDA_IMM, 5)], [W(o.LDA_IMM, 5)], 0x01),
             ([W(o.CPX_M), W(o.CPY_IMM, 0x13)], [W(o.CPY_IMM, 0x12)], 0xFF),
             ([W(o.LDA_IMM, 1), W(o.EOR_IMM, 0xDC)], [W(o.LDA_IMM, 0xDD)], 0),
             ([W(o.SEC), W(o.ROL_A)], [W(o.ASL_A)], 0)]
    for t, c, dead in pairs:
        assert search.quick_check_masked(t, c, dead) == o.quick_check(t, c, dead)
        assert search.mid_check_masked(t, c, dead) == o.mid_check(t, c, dead)
        assert search.exhaustive_check_masked(t, c, dead) == o.exhaustive_check(t, c, dead, mode=2)
        assert search.quick_check(t, c) == o.quick_check(t, c, 0)
        assert search.exhaustive_check(t, c) == o.exhaustive_check(t, c, 0, mode=2)
        assert search.flag_diff(t, c) == o.flag_diff(t, c)


def test_search_single_readme_example():
    W = o.word
    r = search.search_single([W(o.TXA), W(o.TAX)])
    assert r is not None and r.replacement == (W(o.TXA),) and r.bytes_saved == 1 and r.cycles_saved == 2
    assert search.search_single([W(o.LDA_IMM, 1), W(o.LDX_IMM, 2)]) is None
    r = search.search_single([W(o.CLC), W(o.LDA_IMM, 5)], dead_flags=0x01)
    assert r is not None and r.replacement == (W(o.LDA_IMM, 5),) and r.dead_flags == 0x01


def test_writers_match_the_command_line():
    with tempfile.TemporaryDirectory() as d:
        js, jl = os.path.join(d, "r.json"), os.path.join(d, "r.jsonl")
        p = subprocess.run([CLI, "enumerate", "--max-target", "2", "--dead-flags", "0xFF", "--first-target", "0",
                            "--end-target", "5000", "--output", js, "--jsonl", jl],
                           capture_output=True, text=True, timeout=300)
        assert p.returncode == 0, p.stderr
        cli_json, cli_jsonl = open(js).read(), open(jl).read()
    rules = search.run(search.Config(max_target_len=2, dead_flags=0xFF, first_target=0, end_target=5000))
    assert rules and any(r.dead_flags for r in rules)
    f = io.StringIO()
    result.write_json(f, rules)
    assert f.getvalue() == cli_json
    f = io.StringIO()
    result.write_jsonl(f, rules, bytes_model=isa.BYTES_GO)
    assert f.getvalue() == cli_jsonl
