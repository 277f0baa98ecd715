"""Launch-geometry autotuner.

Two knobs shape the kernels (ol_tune_set):

* ``scan_bpt``    bytes per lane of the scan kernel (64, 128, 256): a scan tile
                  is 256 lanes x scan_bpt bytes;
* ``tree_rounds`` rounds of 64 lines per tree wave (1, 4, 16): a tree tile is
                  256 x tree_rounds lines, the per-wave tile 64 x tree_rounds.

The best pair depends on the workload (one large document wants large tiles;
a batch of small documents pads every document to whole tree tiles).  The
tuner times every pair on the caller's own workload with HIP events around
the kernels, keeps the fastest, and remembers it per workload class in a
small JSON file next to this module (``tuning.json``), so later runs start
tuned without sweeping again.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Callable, Dict, Optional

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_PATH = os.path.join(HERE, "tuning.json")

SCAN_BPT = (64, 128, 256)
TREE_ROUNDS = (1, 4, 16)
KERNELS = ("scan_classify", "scan_fixup", "tree", "many", "clear")


def _load() -> Dict[str, dict]:
    try:
        with open(TUNING_PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _save(table: Dict[str, dict]) -> None:
    tmp = TUNING_PATH + ".tmp"
    with open(tmp, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    os.replace(tmp, TUNING_PATH)


def kernel_ms(parser, step: Callable[[], None], reps: int = 3) -> float:
    """Mean kernel milliseconds of one `step` (HIP events on the context's stream)."""
    step()
    parser.synchronize()
    parser.profile_take()
    parser.profile_enable(True)
    for _ in range(reps):
        step()
    parser.synchronize()
    parser.profile_enable(False)
    prof = parser.profile_take()
    return sum(ms for name, (_, ms) in prof.items() if name in KERNELS) / reps


def tune(parser, workload: str, step: Callable[[], None], reps: int = 3,
         scan_bpt=SCAN_BPT, tree_rounds=TREE_ROUNDS, persist: bool = True) -> dict:
    """Times `step` under every (scan_bpt, tree_rounds) pair, applies the
    fastest to `parser` and returns {"scan_bpt", "tree_rounds", "kernel_ms",
    "sweep": {...}}.  `step` must re-run the workload (it is called once per
    pair to size, then `reps` times)."""
    sweep = {}
    best = None
    for bpt, rounds in itertools.product(scan_bpt, tree_rounds):
        parser.tune_set("scan_bpt", bpt)
        parser.tune_set("tree_rounds", rounds)
        ms = kernel_ms(parser, step, reps)
        sweep[f"{bpt}/{rounds}"] = round(ms, 4)
        if best is None or ms < best[0]:
            best = (ms, bpt, rounds)
    ms, bpt, rounds = best
    parser.tune_set("scan_bpt", bpt)
    parser.tune_set("tree_rounds", rounds)
    rec = {"scan_bpt": bpt, "tree_rounds": rounds, "kernel_ms": round(ms, 4), "sweep": sweep}
    if persist:
        table = _load()
        table[workload] = rec
        _save(table)
    return rec


def apply(parser, workload: str) -> Optional[dict]:
    """Applies a remembered choice for `workload` (None if there is none)."""
    rec = _load().get(workload)
    if not rec:
        return None
    parser.tune_set("scan_bpt", int(rec["scan_bpt"]))
    parser.tune_set("tree_rounds", int(rec["tree_rounds"]))
    return rec
