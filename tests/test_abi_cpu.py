"""CPU-side checks of lib6502core: the C ABI loads and exports every symbol
include/sx6502.h declares, host-side entry points validate their inputs, and
the dependency table behind the projected exhaustive sweep is sound.

No compute call needs a GPU here; the sweep-plan audit evaluates the planned
sub-sweeps with the CPU oracle's executor.
"""
import os
import re

import numpy as np
import pytest

import oracle as o
from sx6502 import core, isa

HDR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "sx6502.h")


def test_every_declared_symbol_is_exported():
    decl = set(re.findall(r"^(?:int|uint64_t) (sx_\w+)\(", open(HDR).read(), re.M))
    assert decl == set(core.EXPORTS)
    L = core.lib()
    for name in decl:
        assert hasattr(L, name), name
    assert L.sx_abi_version() == 3


def test_open_without_device_fails_cleanly():
    try:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    except Exception:
        pass
    with pytest.raises(core.SxError) as e:
        core.Device(0)
    assert e.value.code == core.SX_ENODEV


def test_plan_rejects_bad_opcode():
    with pytest.raises(core.SxError):
        core.exh_plan([o.word(58)], [o.word(0)])


def test_full_plan_is_the_go_domain():
    W = o.word
    sets, n = core.exh_plan([W(o.CPX_M), W(o.CPY_M)], [W(o.CPY_M)], 0, core.EXH_FULL)
    assert n == 2 ** 33 and len(sets) == 1          # A, C, X, Y, M
    sets, n = core.exh_plan([W(o.TXA), W(o.TAX)], [W(o.TXA)], 0, core.EXH_FULL)
    assert n == 2 ** 17
    sets, n = core.exh_plan([W(o.CLC), W(o.CLC)], [W(o.CLC)], 0, core.EXH_FULL)
    assert n == 512


BYTES = (0, 1, 2, 4, 5, 6, 7)


def _outputs(st):
    """Outputs in the dependency numbering: bytes, then flag bits 8..15."""
    out = {o: st[..., o] for o in BYTES}
    for b in range(8):
        out[8 + b] = (st[..., 3] >> b) & 1
    return out


def test_dependency_table_is_sound():
    """For every opcode (immediate opcodes at all 256 immediates), output o and
    input q outside dep[op][o], changing q (all 256 values of a byte, both
    values of a flag bit) on random base states never changes output o."""
    dep = core.dep_table()
    rng = np.random.default_rng(5)
    inputs = [(q, 256) for q in BYTES] + [(8 + b, 2) for b in range(8)]
    for op in range(58):
        imms = np.arange(256, dtype=np.uint32) if "_IMM" in o.OPS[op] else np.zeros(1, np.uint32)
        nb = 8 if len(imms) > 1 else 64
        base = rng.integers(0, 256, (nb, 8)).astype(np.uint8)
        for q, nv in inputs:
            # axes (immediate, base state, value of q)
            st = np.broadcast_to(base[None, :, None, :], (len(imms), nb, nv, 8)).copy()
            vals = np.arange(nv, dtype=np.uint8)[None, None, :]
            if q < 8:
                st[..., q] = vals
            else:
                b = q - 8
                st[..., 3] = (st[..., 3] & ~np.uint8(1 << b)) | (vals << b)
            words = np.broadcast_to((op | (imms << 16))[:, None, None], (len(imms), nb, nv)).ravel()
            res = o.exec_many(st.reshape(-1, 8), words).reshape(len(imms) * nb, nv, 8)
            for out_i, col in _outputs(res).items():
                d = int(dep[op, out_i])
                deps = (1 << out_i) if d == 0xFFFF else d
                if deps & (1 << q):
                    continue
                assert (col == col[:, :1]).all(), (o.OPS[op], "out", out_i, "in", q)


def test_flag_projection_shrinks_compare_chains():
    W = o.word
    # CMP #n : CPX #m -> CPX #m: N, Z, C come from X alone
    sets, n = core.exh_plan([W(o.CMP_IMM, 5), W(o.CPX_IMM, 7)], [W(o.CPX_IMM, 7)], 0, core.EXH_PROJECTED)
    assert n <= 2 * 512


def _sweep_ok(t, c, dead, vset):
    """Evaluate one planned sub-sweep with the oracle executor."""
    regs = [r for r in range(8) if vset & (1 << r)]
    sizes = [2 if r == 3 else 256 for r in regs]
    n = int(np.prod(sizes)) if regs else 1
    st = np.zeros((n, 8), np.uint8)
    st[:, 4] = 0xFD
    idx = np.arange(n)
    for r, s in zip(regs, sizes):
        st[:, r] = idx % s
        idx //= s
    ts, cs = st.copy(), st.copy()
    for w in t:
        ts = o.exec_many(ts, np.uint32(w))
    for w in c:
        cs = o.exec_many(cs, np.uint32(w))
    pm = 0xCF & ~dead
    diff = ts != cs
    diff[:, 3] = (ts[:, 3] & pm) != (cs[:, 3] & pm)
    return not diff.any()


def test_projected_plan_decides_like_the_complete_sweep():
    """On rule pairs of the C1 golden set and perturbations of them, the
    projected sub-sweeps accept exactly when the oracle's complete sweep does
    (pairs whose sub-sweeps are at most 2^17 states)."""
    gold = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
    rules = np.concatenate([np.load(os.path.join(gold, "c1_go_deadFF.npy"))[:600]])
    rng = np.random.default_rng(3)
    tab = o.instruction_table(0)
    checked = 0
    for r in rules:
        t = r["src"][: r["src_len"]].tolist()
        c = r["rep"][: r["rep_len"]].tolist()
        dead = int(r["dead_flags"]) and 0xFF
        for cand in (c, [int(tab[rng.integers(0, len(tab))])]):
            sets, total = core.exh_plan(t, cand, dead, core.EXH_PROJECTED)
            if total > (1 << 18):
                continue
            got = all(_sweep_ok(t, cand, dead, s) for s in sets)
            want = o.exhaustive_check(t, cand, dead, mode=1)
            assert got == want, (isa.disasm_seq(t), isa.disasm_seq(cand), dead, sets)
            checked += 1
    assert checked > 500


def test_ctypes_mirrors_match_the_c_header(tmp_path):
    """sizeof/offsetof of every ABI struct, from a C program compiled against
    include/sx6502.h, equal the ctypes/numpy mirrors used by the package."""
    import ctypes as C
    import subprocess
    structs = {"sx_rule": core.RuleC, "sx_enum_cfg": core.EnumCfg, "sx_stats": core.Stats,
               "sx_stoke_cfg": core.StokeCfg, "sx_stoke_result": core.StokeResult,
               "sx_stoke_chain": core.StokeChain, "sx_target_result": core.TargetResult,
               "sx_stoke_found": core.StokeFound}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sx6502.h"', 'int main(void) {']
    for name, py in structs.items():
        lines.append(f'  printf("{name} size %zu\\n", sizeof({name}));')
        for f, _ in py._fields_:
            lines.append(f'  printf("{name}.{f} %zu\\n", offsetof({name}, {f.rstrip("_")}));')
    lines.append("  return 0; }")
    src = tmp_path / "sizes.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "sizes"
    subprocess.run(["gcc", "-I", os.path.dirname(HDR), str(src), "-o", str(exe)], check=True)
    got = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    for name, py in structs.items():
        assert int(got[f"{name} size"]) == C.sizeof(py), name
        for f, _ in py._fields_:
            assert int(got[f"{name}.{f}"]) == getattr(py, f).offset, (name, f)
    assert int(got["sx_enum_cfg.include_empty"]) == 72  # SX_ENUM_CFG_SIZE_V1: the layout before it
    assert core.STOKE_RESULT_DTYPE.itemsize == C.sizeof(core.StokeResult)
    assert core.RULE_DTYPE.itemsize == C.sizeof(core.RuleC)
