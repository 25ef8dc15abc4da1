"""The search API of pkg/search, served by the device.

``Config`` / ``run`` mirror search.Config and search.Run
(pkg/search/search.go:13-54); ``search_single`` mirrors SearchSingle(Masked)
(search.go:78-106); the predicate functions mirror QuickCheck, MidCheck,
ExhaustiveCheck, their *Masked variants and FlagDiff
(pkg/search/verifier.go:84-105,262,427-466,591-604).  All of them run on the
GPU through lib6502core.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Sequence

import numpy as np

from . import core, isa
from .result import Rule, rules_from_array

_dev: Optional[core.Device] = None


def device() -> core.Device:
    global _dev
    if _dev is None:
        _dev = core.Device(0)
    return _dev


@dataclass
class Config:
    max_target_len: int = 2
    max_cand_len: int = 0          # 0 = max_target_len - 1 (search.go:26-28)
    dead_flags: int = isa.DEAD_NONE
    verbose: bool = False
    # device-side semantics switches (see include/sx6502.h)
    bytes_model: int = isa.BYTES_GO
    hits: int = core.HITS_FIRST
    exh_mode: int = core.EXH_PROJECTED
    use_mid: bool = True
    targets: Optional[np.ndarray] = None
    first_target: int = 0
    end_target: int = 0
    stats: dict = field(default_factory=dict)


def run(cfg: Config) -> List[Rule]:
    """search.Run (pkg/search/search.go:21-54): every target length from 2 to
    max_target_len, each against candidates of length 1..max_cand_len
    (max_target_len - 1 when 0, for every target length: search.go:26-28,33);
    given targets (one length), those alone.  Rules in target order, as the
    device emits them (Table.Rules ordering: result.sorted_rules)."""
    max_cand = cfg.max_cand_len or cfg.max_target_len - 1
    if cfg.targets is not None:
        lengths = [cfg.targets.shape[1]]
    else:
        lengths = list(range(2, cfg.max_target_len + 1))
    rules: List[Rule] = []
    for tl in lengths:
        arr, st = device().enumerate(cfg.targets, target_len=tl, max_cand_len=max_cand, dead=cfg.dead_flags,
                                     bytes_model=cfg.bytes_model, hits=cfg.hits, exh_mode=cfg.exh_mode,
                                     use_mid=cfg.use_mid, first=cfg.first_target, end=cfg.end_target)
        cfg.stats[tl] = st
        rules.extend(rules_from_array(arr))
    return rules


def _one(t: Sequence[int], c: Sequence[int], stage: int, dead: int) -> bool:
    return bool(device().check_pairs([list(t)], [list(c)], stage, dead)[0] & stage)


def quick_check(target, candidate) -> bool:
    return _one(target, candidate, core.STAGE_QC, 0)


def quick_check_masked(target, candidate, dead: int) -> bool:
    return _one(target, candidate, core.STAGE_QC, dead)


def mid_check(target, candidate) -> bool:
    return _one(target, candidate, core.STAGE_MID, 0)


def mid_check_masked(target, candidate, dead: int) -> bool:
    return _one(target, candidate, core.STAGE_MID, dead)


def exhaustive_check(target, candidate) -> bool:
    return _one(target, candidate, core.STAGE_EXH, 0)


def exhaustive_check_masked(target, candidate, dead: int) -> bool:
    return _one(target, candidate, core.STAGE_EXH, dead)


def flag_diff(target, candidate) -> int:
    return int(device().flag_diff([list(target)], [list(candidate)])[0])


def search_single(target: Sequence[int], max_cand_len: int = 1, dead_flags: int = 0) -> Optional[Rule]:
    """Shortest replacement of one target (length 2 or 3), first hit in Go
    order; with dead_flags the flag-relaxed rule is returned when the exact
    search finds none, as Table.Rules ordering would put either first."""
    t = np.array([list(target)], np.uint32)
    arr, _ = device().enumerate(t, target_len=len(target), max_cand_len=max_cand_len, dead=dead_flags)
    rules = rules_from_array(arr)
    if not rules:
        return None
    rules.sort(key=lambda r: (-r.bytes_saved, -r.cycles_saved))
    return rules[0]
