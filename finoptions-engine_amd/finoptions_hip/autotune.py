"""Launch-parameter autotuner for the evaluation kernel's strike tile.

The evaluation pass regenerates every path once per strike tile, so the tile trades path
regeneration (fewer, larger tiles) against register pressure (per-strike accumulators). The best
tile depends on strikes per group, observation rows, path counts and whether Greeks ride along.
``autotune`` times each legal tile on a reduced copy of the request and installs the fastest on
the executor. Reduction trees are fixed by 1024-path chunks, so every tile gives bitwise identical
results; only time changes.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable

from .simulation import HipExecutor

PLAIN_TILES = (1, 4, 16)
GREEK_TILES = (1, 4, 6, 8)


@dataclass
class TuningCache:
    """Best tile per (precision, strikes per group, observation rows, greeks) shape key."""

    best: dict[tuple, int] = field(default_factory=dict)
    timings: dict[tuple, dict[int, float]] = field(default_factory=dict)


_CACHE = TuningCache()


def shape_key(executor: HipExecutor, strikes_per_group: int, observation_rows: int,
              greeks: bool) -> tuple:
    return (executor.precision.value, int(strikes_per_group), int(observation_rows), bool(greeks))


def autotune(executor: HipExecutor, run_sample: Callable[[], object], *, strikes_per_group: int,
             observation_rows: int, greeks: bool, repeats: int = 2,
             cache: TuningCache | None = None) -> int:
    """Time ``run_sample`` (a reduced pricing call) under each legal tile; install the fastest.

    ``run_sample`` must price a request of the same shape class (strikes per group, rows, Greeks)
    with fewer paths or groups. Returns the chosen tile.
    """
    cache = cache or _CACHE
    key = shape_key(executor, strikes_per_group, observation_rows, greeks)
    if key in cache.best:
        executor.tune(eval_strike_tile=cache.best[key])
        return cache.best[key]
    candidates = GREEK_TILES if greeks else PLAIN_TILES
    if strikes_per_group <= 1:
        candidates = (1,)
    timings: dict[int, float] = {}
    for tile in candidates:
        executor.tune(eval_strike_tile=tile)
        run_sample()                      # warm the code object and buffers
        best = float("inf")
        for _ in range(repeats):
            t0 = time.perf_counter()
            run_sample()
            best = min(best, time.perf_counter() - t0)
        timings[tile] = best
    choice = min(timings, key=timings.get)
    cache.best[key] = choice
    cache.timings[key] = timings
    executor.tune(eval_strike_tile=choice)
    return choice


__all__ = ["GREEK_TILES", "PLAIN_TILES", "TuningCache", "autotune", "shape_key"]
