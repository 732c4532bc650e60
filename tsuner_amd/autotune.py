"""Tile-height autotuner for the stage kernels.

The stage kernel's tile is 60 columns (one wave) by ``tile_rows`` rows; the
best height trades vertical-halo overhead (short tiles) against too few
waves to fill 256 CUs (tall tiles on small grids).  The tuner times
whole-ensemble steps with HIP events for each candidate height, one grid-size
class at a time (coordinate descent, the other classes held fixed), restores
the ensemble state after each trial, and caches the choice per
(GPU architecture, precision, grid-size class) in a JSON file.
"""
from __future__ import annotations

import json
import os
from typing import Dict, Sequence

CANDIDATES = (4, 8, 16, 32, 64)


def cache_path() -> str:
    base = os.environ.get("XDG_CACHE_HOME") or os.path.join(os.path.expanduser("~"), ".cache")
    return os.path.join(base, "tsuner_amd", "tile_rows.json")


def _load() -> Dict[str, int]:
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(d: Dict[str, int]) -> None:
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        with open(p, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
    except OSError:
        pass


def size_class(npx: int, npy: int) -> int:
    return max(npx, npy)


def key(arch: str, precision: str, cls: int) -> str:
    return "%s/%s/%d" % (arch, precision, cls)


def tune_ensemble(ens, precision: str, arch: str = "gfx950", candidates: Sequence[int] = CANDIDATES,
                  steps: int = 3, use_cache: bool = True, log=None) -> Dict[int, int]:
    """Chooses tile rows per size class for an Ensemble; returns {class: rows}
    and applies them."""
    classes = sorted({size_class(nx, ny) for ny, nx in ens.shapes})
    cache = _load() if use_cache else {}
    choice = {c: cache.get(key(arch, precision, c), 16) for c in classes}
    todo = [c for c in classes if key(arch, precision, c) not in cache]

    def apply():
        ens.set_tile_rows([choice[size_class(nx, ny)] for ny, nx in ens.shapes])

    if todo:
        ens.snapshot(False)
        apply()
        ens.time_steps(1)   # warm-up
        ens.snapshot(True)
        for c in todo:
            best, best_ms = choice[c], None
            for r in candidates:
                choice[c] = r
                apply()
                ms = ens.time_steps(steps)
                ens.snapshot(True)
                if log:
                    log("autotune %s class %d rows %d: %.3f ms/step" % (precision, c, r, ms / steps))
                if best_ms is None or ms < best_ms:
                    best, best_ms = r, ms
            choice[c] = best
            cache[key(arch, precision, c)] = best
        if use_cache:
            _store(cache)
    apply()
    return choice
