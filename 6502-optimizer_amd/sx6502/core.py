"""ctypes binding of lib6502core.so (include/sx6502.h).

This is the only way the package reaches the device: every search, check and
execution call goes through the C ABI.  Loading fails loudly when the library
has not been built; there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib6502core.so")

SX_OK, SX_EINVAL, SX_EOPCODE, SX_ECAPACITY, SX_EHIP, SX_ENOMEM, SX_ENODEV = 0, -1, -2, -3, -4, -5, -6
STAGE_QC, STAGE_MID, STAGE_EXH = 1, 2, 4
HITS_FIRST, HITS_ALL = 0, 1
EXH_PROJECTED, EXH_FULL = 0, 1

# Every symbol include/sx6502.h declares.
EXPORTS = ("sx_abi_version", "sx_open", "sx_close", "sx_last_error", "sx_enumerate", "sx_fetch_rules",
           "sx_pruned_targets", "sx_check_pairs", "sx_flag_diff", "sx_exec_states", "sx_exh_plan",
           "sx_dep_table", "sx_stoke_run", "sx_stoke_collect", "sx_mlp_param_count", "sx_mlp_create", "sx_mlp_destroy",
           "sx_mlp_get_params", "sx_mlp_set_params", "sx_mlp_loss_grad", "sx_mlp_train", "sx_mlp_predict",
           "sx_proposal_weights", "sx_mlp_proposal", "sx_mfma_gemm_test", "sx_target_search",
           "sx_measure_int32_rate", "sx_check_pairs_noalloc", "sx_fetch_empty_rules")
NOALLOC_MAX_PAIRS = 65536
MLP_CLASSES = 2863   # instructions
MLP_OUTPUTS = 2871   # instructions, then 8 context-position pointers
MLP_CTX = 8
PAD_TOKEN = 2863

class RuleC(C.Structure):
    _fields_ = [("src", C.c_uint32 * 4), ("rep", C.c_uint32 * 4), ("target_index", C.c_uint64),
                ("src_len", C.c_uint8), ("rep_len", C.c_uint8), ("dead_flags", C.c_uint8), ("pass_", C.c_uint8),
                ("bytes_saved", C.c_int16), ("cycles_saved", C.c_int16)]


RULE_DTYPE = np.dtype([
    ("src", "<u4", (4,)), ("rep", "<u4", (4,)), ("target_index", "<u8"),
    ("src_len", "u1"), ("rep_len", "u1"), ("dead_flags", "u1"), ("pass_", "u1"),
    ("bytes_saved", "<i2"), ("cycles_saved", "<i2"),
])
assert RULE_DTYPE.itemsize == C.sizeof(RuleC) == 48


class EnumCfg(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("target_len", C.c_int32), ("max_cand_len", C.c_int32),
                ("bytes_model", C.c_int32), ("hits", C.c_int32), ("exh_mode", C.c_int32),
                ("use_mid", C.c_int32), ("dead_flags", C.c_uint32), ("targets", C.POINTER(C.c_uint32)),
                ("n_targets", C.c_uint64), ("first_target", C.c_uint64), ("end_target", C.c_uint64),
                ("tile_targets", C.c_uint32), ("survivor_cap", C.c_uint32), ("include_empty", C.c_int32),
                ("reserved", C.c_uint32)]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in ("targets", "qc_pairs", "qc_raw_hits", "qc_hits", "mid_pass",
                                          "exh_pairs", "exh_states", "exh_rounds", "rules",
                                          "overflow_retries")] + \
               [(n, C.c_double) for n in ("ms_setup", "ms_qc", "ms_mid", "ms_exh", "ms_total",
                                          "ms_qc_kernel", "ms_exh_kernel")] + \
               [(n, C.c_uint64) for n in ("qc_launches", "exh_launches")] + \
               [(n, C.c_uint32) for n in ("qc_tile", "qc_blocks_per_cu")] + [("empty_rules", C.c_uint64)] + \
               [("ms_exh_group_kernel", C.c_double)] + \
               [(n, C.c_uint64) for n in ("exh_group_launches", "exh_group_jobs", "exh_group_states")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class StokeCfg(C.Structure):
    _fields_ = [("struct_size", C.c_uint32), ("chains_per_target", C.c_uint32), ("steps", C.c_uint64),
                ("steps_per_launch", C.c_uint32), ("stop_when_verified", C.c_int32), ("t0", C.c_double),
                ("decay", C.c_double), ("seed", C.c_uint64), ("dead_flags", C.c_uint32), ("pad", C.c_uint32),
                ("weights", C.POINTER(C.c_uint32))]


class StokeResult(C.Structure):
    _fields_ = [("rep", C.c_uint32 * 10), ("rep_len", C.c_uint8), ("verified", C.c_uint8), ("dead_flags", C.c_uint8),
                ("pad", C.c_uint8),
                ("bytes_saved", C.c_int16), ("cycles_saved", C.c_int16), ("chain", C.c_uint32),
                ("step", C.c_uint64), ("seconds", C.c_double)]


class StokeFound(C.Structure):
    _fields_ = [("target", C.c_uint32), ("chain", C.c_uint32), ("step", C.c_uint64), ("seconds", C.c_double),
                ("rep", C.c_uint32 * 10), ("rep_len", C.c_uint8), ("dead_flags", C.c_uint8),
                ("bytes_saved", C.c_int16), ("cycles_saved", C.c_int16), ("pad", C.c_uint16)]


class StokeChain(C.Structure):
    _fields_ = [("seq", C.c_uint32 * 10), ("len", C.c_int32), ("cost10", C.c_int32), ("step", C.c_uint64),
                ("accepted", C.c_uint64)]


# numpy views with the C layout (alignment included), checked against ctypes
STOKE_RESULT_DTYPE = np.dtype(StokeResult)
STOKE_CHAIN_DTYPE = np.dtype(StokeChain)
STOKE_FOUND_DTYPE = np.dtype(StokeFound)
assert STOKE_FOUND_DTYPE.itemsize == C.sizeof(StokeFound) == 72
assert STOKE_RESULT_DTYPE.itemsize == C.sizeof(StokeResult) == 72
assert STOKE_CHAIN_DTYPE.itemsize == C.sizeof(StokeChain) == 64


class TargetResult(C.Structure):
    _fields_ = [("rep", C.c_uint32 * 4), ("rep_len", C.c_uint8), ("found", C.c_uint8), ("pass_", C.c_uint8),
                ("dead_flags", C.c_uint8), ("bytes_saved", C.c_int16), ("cycles_saved", C.c_int16)]


assert C.sizeof(TargetResult) == 24


class SxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lib6502core error {code}: {msg}")
        self.code = code


_lib = None


def lib() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `make -C {os.path.dirname(_HERE)}` "
                           "(python __graft_entry__.py build)")
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    u32p, u8p = C.POINTER(C.c_uint32), C.POINTER(C.c_uint8)
    L.sx_abi_version.restype = C.c_int
    L.sx_open.argtypes = [C.c_int, C.POINTER(vp)]
    L.sx_close.argtypes = [vp]
    L.sx_last_error.argtypes = [vp, C.c_char_p, C.c_size_t]
    L.sx_enumerate.argtypes = [vp, C.POINTER(EnumCfg), vp, C.c_uint64, C.POINTER(C.c_uint64),
                               C.POINTER(Stats)]
    L.sx_fetch_rules.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
    L.sx_fetch_empty_rules.argtypes = [vp, vp, C.c_uint64, C.c_uint64, C.POINTER(C.c_uint64)]
    L.sx_pruned_targets.argtypes = [vp, C.c_uint64, C.c_uint64, u32p, C.POINTER(C.c_uint64)]
    L.sx_check_pairs.argtypes = [vp, u32p, u8p, u32p, u8p, C.c_uint64, C.c_uint32, C.c_uint32,
                                 C.c_int32, u8p]
    L.sx_check_pairs_noalloc.argtypes = L.sx_check_pairs.argtypes
    L.sx_flag_diff.argtypes = [vp, u32p, u8p, u32p, u8p, C.c_uint64, u8p]
    L.sx_exec_states.argtypes = [vp, u8p, u32p, u8p, C.c_uint64]
    L.sx_exh_plan.argtypes = [u32p, C.c_int, u32p, C.c_int, C.c_uint32, C.c_int32, u8p,
                              C.POINTER(C.c_int32), C.POINTER(C.c_uint64)]
    L.sx_dep_table.argtypes = [C.POINTER(C.c_uint16)]
    L.sx_stoke_run.argtypes = [vp, C.POINTER(StokeCfg), u32p, u8p, C.c_uint64, vp, vp]
    L.sx_stoke_collect.argtypes = [vp, C.POINTER(StokeCfg), u32p, u8p, C.c_uint64, vp, C.c_uint64,
                                   C.POINTER(C.c_uint64), vp]
    u16p, f32p = C.POINTER(C.c_uint16), C.POINTER(C.c_float)
    L.sx_mlp_create.argtypes = [vp, C.c_uint64, C.POINTER(vp)]
    L.sx_mlp_destroy.argtypes = [vp]
    L.sx_mlp_get_params.argtypes = [vp, f32p]
    L.sx_mlp_set_params.argtypes = [vp, f32p]
    L.sx_mlp_loss_grad.argtypes = [vp, u16p, u16p, C.c_uint32, f32p, f32p]
    L.sx_mlp_train.argtypes = [vp, u16p, u16p, C.c_uint32, C.c_uint32, C.c_uint32, C.c_float, C.c_uint64, f32p]
    L.sx_mlp_predict.argtypes = [vp, u16p, C.c_uint32, f32p]
    L.sx_proposal_weights.argtypes = [f32p, C.c_uint64, C.c_double, u32p]
    L.sx_mlp_proposal.argtypes = [vp, u16p, C.c_uint32, C.c_double, u32p]
    L.sx_mfma_gemm_test.argtypes = [vp, u16p, u16p, C.c_int, C.c_int, C.c_int, f32p]
    L.sx_measure_int32_rate.argtypes = [vp, C.POINTER(C.c_double)]
    L.sx_target_search.argtypes = [vp, u32p, C.c_int, C.c_int, C.c_uint32, C.c_int, C.POINTER(TargetResult),
                                   C.POINTER(C.c_uint64)]
    for name in EXPORTS:
        getattr(L, name).restype = C.c_int
    L.sx_mlp_param_count.restype = C.c_uint64
    _lib = L
    return L


def _p32(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint32))


def _p8(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint8))


def _p16(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_uint16))


def _pf(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def proposal_weights(probs: np.ndarray, eps: float = 0.1) -> np.ndarray:
    """sx_proposal_weights: n x 2,863 probabilities -> uint32 tables summing
    to 2^32 with a uniform share eps (host-side, no device needed)."""
    p = np.ascontiguousarray(probs, dtype=np.float32)
    assert p.ndim == 2 and p.shape[1] == MLP_CLASSES
    out = np.zeros(p.shape, np.uint32)
    rc = lib().sx_proposal_weights(_pf(p), len(p), float(eps), _p32(out))
    if rc != SX_OK:
        raise SxError(rc, "sx_proposal_weights: bad probabilities or eps")
    return out


def pack_seqs(seqs: Sequence[Sequence[int]]):
    """Sequences of up to 4 words -> (n, 4) uint32 words and (n,) lengths."""
    n = len(seqs)
    w = np.zeros((n, 4), np.uint32)
    ln = np.zeros(n, np.uint8)
    for i, s in enumerate(seqs):
        if len(s) > 4:
            raise ValueError("sequence longer than 4")
        w[i, :len(s)] = s
        ln[i] = len(s)
    return w, ln


def exh_plan(t: Sequence[int], c: Sequence[int], dead: int = 0, mode: int = EXH_PROJECTED):
    """Sweep plan of one pair: (list of varying-register sets, total states)."""
    ta = np.array(list(t) or [0], np.uint32)
    ca = np.array(list(c) or [0], np.uint32)
    sets = np.zeros(8, np.uint8)
    ns = C.c_int32(0)
    st = C.c_uint64(0)
    rc = lib().sx_exh_plan(_p32(ta), len(t), _p32(ca), len(c), dead, mode, _p8(sets), C.byref(ns), C.byref(st))
    if rc:
        raise SxError(rc, "sx_exh_plan")
    return [int(x) for x in sets[:ns.value]], st.value


def dep_table() -> np.ndarray:
    """(58, 16) dependency table of the projected sweep: inputs/outputs A=0,
    X=1, Y=2, S=4, M=5, S0=6, S1=7 and 8 + b for flag bit b; 0xFFFF = the
    output is not written."""
    out = np.zeros(58 * 16, np.uint16)
    rc = lib().sx_dep_table(out.ctypes.data_as(C.POINTER(C.c_uint16)))
    if rc:
        raise SxError(rc, "sx_dep_table")
    return out.reshape(58, 16)


class Device:
    """One lib6502core handle on one HIP device."""

    def __init__(self, device: int = 0):
        self._h = C.c_void_p()
        rc = lib().sx_open(device, C.byref(self._h))
        if rc:
            raise SxError(rc, f"sx_open({device}) failed (no usable gfx950 device?)")

    def close(self):
        if self._h:
            lib().sx_close(self._h)
            self._h = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc: int, what: str):
        if rc:
            buf = C.create_string_buffer(512)
            lib().sx_last_error(self._h, buf, 512)
            raise SxError(rc, f"{what}: {buf.value.decode()}")

    # -- enumeration ----------------------------------------------------
    def enumerate(self, targets: Optional[np.ndarray] = None, *, target_len: int = 2, max_cand_len: int = 1,
                  dead: int = 0, bytes_model: int = 0, hits: int = HITS_FIRST, exh_mode: int = EXH_PROJECTED,
                  use_mid: bool = True, first: int = 0, end: int = 0, tile: int = 0, survivor_cap: int = 0,
                  include_empty: bool = False, out: Optional[np.ndarray] = None):
        """sx_enumerate: (rules, stats).  With include_empty the empty
        candidate's stream is read with fetch_empty_rules().  out: a rule
        array (RULE_DTYPE) to fetch into when it is large enough (the result
        is then a view of its first records); else a new array."""
        cfg = EnumCfg()
        cfg.struct_size = C.sizeof(EnumCfg)
        cfg.target_len = target_len
        cfg.max_cand_len = max_cand_len
        cfg.bytes_model = bytes_model
        cfg.hits = hits
        cfg.exh_mode = exh_mode
        cfg.use_mid = 1 if use_mid else 0
        cfg.dead_flags = dead
        keep = None
        if targets is not None:
            keep = np.ascontiguousarray(targets, dtype=np.uint32)
            if keep.ndim != 2 or keep.shape[1] != target_len:
                raise ValueError("targets must have shape (n, target_len)")
            cfg.targets = _p32(keep)
            cfg.n_targets = keep.shape[0]
        cfg.first_target, cfg.end_target = first, end
        cfg.tile_targets, cfg.survivor_cap = tile, survivor_cap
        cfg.include_empty = 1 if include_empty else 0
        n = C.c_uint64(0)
        st = Stats()
        rc = lib().sx_enumerate(self._h, C.byref(cfg), None, 0, C.byref(n), C.byref(st))
        self._check(rc, "sx_enumerate")
        if out is None or out.dtype != RULE_DTYPE or out.ndim != 1 or len(out) < n.value or \
                not out.flags.c_contiguous or not out.flags.writeable:
            out = np.empty(n.value, RULE_DTYPE)  # sx_fetch_rules writes every byte of every record it returns
        got = C.c_uint64(0)
        rc = lib().sx_fetch_rules(self._h, out.ctypes.data_as(C.c_void_p), 0, n.value, C.byref(got))
        self._check(rc, "sx_fetch_rules")
        del keep
        return out[: got.value], st.as_dict()

    def fetch_rules(self, offset: int, n: int) -> np.ndarray:
        """sx_fetch_rules: the last enumeration's rules [offset, offset + n)
        (fewer at the end of the list)."""
        out = np.empty(n, RULE_DTYPE)
        got = C.c_uint64(0)
        rc = lib().sx_fetch_rules(self._h, out.ctypes.data_as(C.c_void_p), offset, n, C.byref(got))
        self._check(rc, "sx_fetch_rules")
        return out[: got.value]

    def fetch_empty_rules(self, n: int) -> np.ndarray:
        """The empty-candidate stream of the last enumeration (its size is
        stats["empty_rules"])."""
        out = np.empty(n, RULE_DTYPE)
        got = C.c_uint64(0)
        rc = lib().sx_fetch_empty_rules(self._h, out.ctypes.data_as(C.c_void_p), 0, n, C.byref(got))
        self._check(rc, "sx_fetch_empty_rules")
        return out[: got.value]

    def pruned_targets(self, first: int = 0, n: Optional[int] = None) -> np.ndarray:
        total = C.c_uint64(0)
        self._check(lib().sx_pruned_targets(self._h, 0, 0, None, C.byref(total)), "sx_pruned_targets")
        if n is None:
            n = total.value - first
        out = np.zeros((n, 2), np.uint32)
        self._check(lib().sx_pruned_targets(self._h, first, n, _p32(out), C.byref(total)), "sx_pruned_targets")
        return out

    # -- bulk checks ------------------------------------------------------
    def check_pairs(self, targets, cands, stages: int = STAGE_QC | STAGE_MID | STAGE_EXH, dead: int = 0,
                    exh_mode: int = EXH_PROJECTED) -> np.ndarray:
        tw, tl = pack_seqs(targets) if not isinstance(targets, tuple) else targets
        cw, cl = pack_seqs(cands) if not isinstance(cands, tuple) else cands
        n = len(tl)
        out = np.zeros(n, np.uint8)
        rc = lib().sx_check_pairs(self._h, _p32(tw), _p8(tl), _p32(cw), _p8(cl), n, stages, dead, exh_mode,
                                  _p8(out))
        self._check(rc, "sx_check_pairs")
        return out

    def check_pairs_noalloc(self, targets, cands, stages: int = STAGE_QC | STAGE_MID | STAGE_EXH, dead: int = 0,
                            exh_mode: int = EXH_PROJECTED) -> np.ndarray:
        """check_pairs through sx_check_pairs_noalloc (no allocation in the
        library; at most NOALLOC_MAX_PAIRS pairs)."""
        tw, tl = pack_seqs(targets) if not isinstance(targets, tuple) else targets
        cw, cl = pack_seqs(cands) if not isinstance(cands, tuple) else cands
        tw, cw = np.ascontiguousarray(tw, np.uint32), np.ascontiguousarray(cw, np.uint32)
        tl, cl = np.ascontiguousarray(tl, np.uint8), np.ascontiguousarray(cl, np.uint8)
        n = len(tl)
        out = np.zeros(n, np.uint8)
        rc = lib().sx_check_pairs_noalloc(self._h, _p32(tw), _p8(tl), _p32(cw), _p8(cl), n, stages, dead,
                                          exh_mode, _p8(out))
        self._check(rc, "sx_check_pairs_noalloc")
        return out

    def flag_diff(self, targets, cands) -> np.ndarray:
        tw, tl = pack_seqs(targets)
        cw, cl = pack_seqs(cands)
        out = np.zeros(len(tl), np.uint8)
        self._check(lib().sx_flag_diff(self._h, _p32(tw), _p8(tl), _p32(cw), _p8(cl), len(tl), _p8(out)),
                    "sx_flag_diff")
        return out

    def exec_states(self, states: np.ndarray, seqs) -> np.ndarray:
        st = np.ascontiguousarray(states, dtype=np.uint8).copy()
        w, ln = pack_seqs(seqs) if not isinstance(seqs, tuple) else seqs
        self._check(lib().sx_exec_states(self._h, _p8(st), _p32(w), _p8(ln), len(ln)), "sx_exec_states")
        return st

    # -- STOKE -----------------------------------------------------------
    @staticmethod
    def _stoke_cfg(targets, chains, steps, steps_per_launch, t0, decay, seed, dead, weights, stop_when_verified):
        n = len(targets)
        tw = np.zeros((n, 8), np.uint32)
        tl = np.zeros(n, np.uint8)
        for i, t in enumerate(targets):
            tw[i, :len(t)] = t
            tl[i] = len(t)
        cfg = StokeCfg()
        cfg.struct_size = C.sizeof(StokeCfg)
        cfg.chains_per_target, cfg.steps, cfg.steps_per_launch = chains, steps, steps_per_launch
        cfg.stop_when_verified = 1 if stop_when_verified else 0
        cfg.t0, cfg.decay, cfg.seed, cfg.dead_flags = t0, decay, seed, dead
        keep = None
        if weights is not None:
            keep = np.ascontiguousarray(weights, dtype=np.uint32)
            assert keep.shape == (n, 2863)
            cfg.weights = _p32(keep)
        return cfg, tw, tl, keep

    def stoke(self, targets, *, chains: int = 16, steps: int = 100000, steps_per_launch: int = 0,
              t0: float = 1.0, decay: float = 0.9999, seed: int = 0, dead: int = 0, weights=None,
              stop_when_verified: bool = True, return_chains: bool = False):
        """STOKE over target sequences (lists of up to 8 words); returns the
        per-target results (STOKE_RESULT_DTYPE) and, on request, the final
        chain states."""
        n = len(targets)
        cfg, tw, tl, keep = self._stoke_cfg(targets, chains, steps, steps_per_launch, t0, decay, seed, dead, weights,
                                            stop_when_verified)
        out = np.zeros(n, STOKE_RESULT_DTYPE)
        ch = np.zeros(n * chains, STOKE_CHAIN_DTYPE) if return_chains else None
        rc = lib().sx_stoke_run(self._h, C.byref(cfg), _p32(tw), _p8(tl), n, out.ctypes.data_as(C.c_void_p),
                                ch.ctypes.data_as(C.c_void_p) if ch is not None else None)
        self._check(rc, "sx_stoke_run")
        del keep
        return (out, ch) if return_chains else out

    def stoke_collect(self, targets, *, chains: int = 16, steps: int = 100000, steps_per_launch: int = 0,
                      t0: float = 1.0, decay: float = 0.9999, seed: int = 0, dead: int = 0, weights=None,
                      cap: int = 1 << 16, return_chains: bool = False):
        """sx_stoke_collect (stoke.Run + Deduplicate): every verified rewrite,
        once per (target, replacement), as STOKE_FOUND_DTYPE, and the number
        found (which exceeds the array's length when cap was too small)."""
        n = len(targets)
        cfg, tw, tl, keep = self._stoke_cfg(targets, chains, steps, steps_per_launch, t0, decay, seed, dead, weights,
                                            False)
        out = np.zeros(cap, STOKE_FOUND_DTYPE)
        ch = np.zeros(n * chains, STOKE_CHAIN_DTYPE) if return_chains else None
        nf = C.c_uint64(0)
        rc = lib().sx_stoke_collect(self._h, C.byref(cfg), _p32(tw), _p8(tl), n, out.ctypes.data_as(C.c_void_p), cap,
                                    C.byref(nf), ch.ctypes.data_as(C.c_void_p) if ch is not None else None)
        self._check(rc, "sx_stoke_collect")
        del keep
        found = out[: min(cap, nf.value)]
        return (found, nf.value, ch) if return_chains else (found, nf.value)

    # -- target command --------------------------------------------------
    def target_search(self, target: Sequence[int], max_cand_len: int = 4, dead: int = 0, bytes_model: int = 0):
        """sx_target_search: ([exact-pass result, masked-pass result], candidates
        scanned); a result is None where its pass found nothing."""
        t = np.ascontiguousarray(list(target), dtype=np.uint32)
        out = (TargetResult * 2)()
        scanned = C.c_uint64(0)
        self._check(lib().sx_target_search(self._h, _p32(t), len(t), max_cand_len, dead, bytes_model, out,
                                           C.byref(scanned)), "sx_target_search")
        res = []
        for r in out:
            res.append(None if not r.found else {
                "rep": [int(r.rep[k]) for k in range(r.rep_len)], "pass": int(r.pass_),
                "dead_flags": int(r.dead_flags), "bytes_saved": int(r.bytes_saved),
                "cycles_saved": int(r.cycles_saved)})
        return res, int(scanned.value)

    # -- learned proposal network ----------------------------------------
    def mlp(self, seed: int = 0) -> "Mlp":
        return Mlp(self, seed)

    def mfma_gemm_test(self, A: np.ndarray, B: np.ndarray) -> np.ndarray:
        """C = A B through the MLP tile kernel; A, B given as uint16 bf16 bits."""
        a = np.ascontiguousarray(A, dtype=np.uint16)
        b = np.ascontiguousarray(B, dtype=np.uint16)
        M, K = a.shape
        K2, N = b.shape
        assert K == K2
        c = np.zeros((M, N), np.float32)
        self._check(lib().sx_mfma_gemm_test(self._h, _p16(a), _p16(b), M, N, K, _pf(c)), "sx_mfma_gemm_test")
        return c

    def measure_int32_rate(self) -> float:
        """The device's int32 VALU rate, lane-operations per second."""
        r = C.c_double(0)
        self._check(lib().sx_measure_int32_rate(self._h, C.byref(r)), "sx_measure_int32_rate")
        return r.value


class Mlp:
    """The learned STOKE proposal network on one device (sx_mlp_* in
    include/sx6502.h).  Contexts are n x 8 uint16 tokens (Go instruction
    index, 2863 = padding), labels n uint16 instruction indices."""

    def __init__(self, dev: Device, seed: int = 0):
        self.dev = dev
        self._m = C.c_void_p()
        dev._check(lib().sx_mlp_create(dev._h, seed, C.byref(self._m)), "sx_mlp_create")
        self.n_params = int(lib().sx_mlp_param_count())

    def close(self):
        if self._m:
            lib().sx_mlp_destroy(self._m)
            self._m = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _ctx(ctx) -> np.ndarray:
        c = np.ascontiguousarray(ctx, dtype=np.uint16)
        assert c.ndim == 2 and c.shape[1] == MLP_CTX
        return c

    def params(self) -> np.ndarray:
        out = np.zeros(self.n_params, np.float32)
        self.dev._check(lib().sx_mlp_get_params(self._m, _pf(out)), "sx_mlp_get_params")
        return out

    def set_params(self, p: np.ndarray):
        a = np.ascontiguousarray(p, dtype=np.float32)
        assert a.shape == (self.n_params,)
        self.dev._check(lib().sx_mlp_set_params(self._m, _pf(a)), "sx_mlp_set_params")

    def loss_grad(self, ctx, labels):
        c = self._ctx(ctx)
        y = np.ascontiguousarray(labels, dtype=np.uint16)
        loss = np.zeros(1, np.float32)
        g = np.zeros(self.n_params, np.float32)
        self.dev._check(lib().sx_mlp_loss_grad(self._m, _p16(c), _p16(y), len(c), _pf(loss), _pf(g)),
                        "sx_mlp_loss_grad")
        return float(loss[0]), g

    def train(self, ctx, labels, steps: int, batch: int = 256, lr: float = 1e-3, seed: int = 0) -> np.ndarray:
        c = self._ctx(ctx)
        y = np.ascontiguousarray(labels, dtype=np.uint16)
        losses = np.zeros(max(steps, 1), np.float32)
        self.dev._check(lib().sx_mlp_train(self._m, _p16(c), _p16(y), len(c), steps, batch, lr, seed,
                                           _pf(losses)), "sx_mlp_train")
        return losses[:steps]

    def predict(self, ctx) -> np.ndarray:
        """n x 2,871 softmax outputs (instructions, then context pointers)."""
        c = self._ctx(ctx)
        out = np.zeros((len(c), MLP_OUTPUTS), np.float32)
        self.dev._check(lib().sx_mlp_predict(self._m, _p16(c), len(c), _pf(out)), "sx_mlp_predict")
        return out

    def proposal(self, ctx, eps: float = 0.1) -> np.ndarray:
        """sx_mlp_proposal: per-context uint32 instruction tables (rows sum
        to 2^32) with the pointer outputs folded onto their instructions."""
        c = self._ctx(ctx)
        out = np.zeros((len(c), MLP_CLASSES), np.uint32)
        self.dev._check(lib().sx_mlp_proposal(self._m, _p16(c), len(c), float(eps), _p32(out)), "sx_mlp_proposal")
        return out
