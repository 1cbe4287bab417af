"""Launch-parameter autotuner for batched LSM requests.

The knobs are the executor's :meth:`HipExecutor.tune` parameters that results are bitwise
invariant to and that a reduced copy of a request measures faithfully:

* ``eval_strike_tile`` -- strikes per evaluation launch column. A group's evaluation paths are
  regenerated (or their stored normals re-read) once per tile, so larger tiles trade path work
  against per-strike accumulators in registers;
* ``fit_cashflows`` -- the fused backward induction's cashflow encoding (binary64 value columns,
  or one stopping byte per strike and path rebuilt from the spot matrix).

Reduction trees are fixed by the path chunks, so every candidate must give bitwise the same
prices, errors and Greeks; the tuner checks that on its sample and raises ``ExecutionError`` if a
candidate does not. The batch schedule and the batch memory are memory-policy choices, left to the
caller.

``tune_request`` prices the request's first groups (about a twentieth of its path work, capped)
under each candidate -- tiles first with value columns, then the stopping rows with the best tile
-- and installs the fastest on the executor. The choice is cached per request shape (method,
Greeks, mesh and dividend kinds, strikes per group, and the observation rows and path counts
rounded up to powers of two). An executor created with ``autotune=True`` runs it on the first
request of every new shape from both batch entry points.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Callable

import numpy as np

from ._lib import ExecutionError

GREEK_TILES = (0, 1, 4, 6, 8)     # 0: the engine's choice by shape
PLAIN_TILES = (0, 1, 4, 16)
SAMPLE_FRACTION = 20              # the sample holds about 1/20 of the request's path work
SAMPLE_WORK_CAP = 4e9             # ... and at most this many strike-path-steps
REPEATS = 2                       # timed runs per candidate (after one warm-up)


@dataclass
class TuningRecord:
    """The tuning chosen for one request shape and the sample timings it came from."""

    key: tuple
    choice: dict
    timings: dict = field(default_factory=dict)   # candidate (tile, cashflows) -> best seconds
    sample_groups: int = 0


@dataclass
class TuningCache:
    records: dict = field(default_factory=dict)
    tuned: int = 0          # samples timed (cache misses)


_CACHE = TuningCache()


def cache() -> TuningCache:
    """The process-wide tuning cache (shared by every executor of one precision)."""
    return _CACHE


def _pow2(n: int) -> int:
    return 1 << max(0, int(n) - 1).bit_length()


def _group_rows(mesh, dividends, groups: int) -> np.ndarray:
    """Observation rows per group (exercise dates + expiry; cash dates of a uniform mesh not
    counted), enough to size the sample and key the shape."""
    from .batch_lsm import ExplicitExerciseMesh, UniformExerciseMesh

    if type(mesh) is ExplicitExerciseMesh:
        return np.diff(mesh.schedules.offsets).astype(np.int64) + 1
    if type(mesh) is UniformExerciseMesh:
        return np.full(groups, mesh.base_intervals, dtype=np.int64)
    offs = getattr(getattr(dividends, "schedules", None), "offsets", None)
    return (np.diff(offs).astype(np.int64) + 1) if offs is not None else np.ones(groups, np.int64)


def shape_key(executor, group_offsets, mesh, dividends, method, greeks, path_counts) -> tuple:
    groups = int(group_offsets.size - 1)
    strikes = int(np.max(np.diff(group_offsets.astype(np.int64)))) if groups else 0
    rows = int(np.max(_group_rows(mesh, dividends, groups))) if groups else 0
    paths = int(max(path_counts.training.max(), path_counts.evaluation.max())) if path_counts is not None \
        else max(method.training_paths, method.evaluation_paths)
    return (executor.precision.value, tuple(sorted(g.value for g in greeks)), type(mesh).__name__,
            type(dividends).__name__, strikes, _pow2(rows), _pow2(paths))


def _slice(n: int, group_offsets, spots, strikes, maturities, rates, vols, dividends, mesh, path_counts):
    """The request's first n groups, with the stream domains they have in the whole request."""
    from .batch_lsm import (CashDividendBatch, ContinuousDividendYieldBatch, EscrowedCashDividendBatch,
                            ExplicitExerciseMesh, RaggedExerciseTimes, RaggedPathCounts)
    from .dividends import RaggedCashDividends

    go = np.ascontiguousarray(group_offsets[:n + 1])
    rows = int(go[-1])

    def ragged_cash(sch):
        off = np.ascontiguousarray(sch.offsets[:n + 1])
        span = int(off[-1] - off[0])
        return RaggedCashDividends(offsets=off, times=np.ascontiguousarray(sch.times[:span]),
                                   amounts=np.ascontiguousarray(sch.amounts[:span]))

    if type(dividends) is ContinuousDividendYieldBatch:
        div = ContinuousDividendYieldBatch(annualized_decimals=np.ascontiguousarray(dividends.annualized_decimals[:n]))
    elif type(dividends) is CashDividendBatch:
        div = CashDividendBatch(schedules=ragged_cash(dividends.schedules), policy=dividends.policy,
                                maturity_boundary=dividends.maturity_boundary)
    else:
        assert type(dividends) is EscrowedCashDividendBatch
        div = EscrowedCashDividendBatch(schedules=ragged_cash(dividends.schedules),
                                        volatility_convention=dividends.volatility_convention,
                                        maturity_boundary=dividends.maturity_boundary)
    if type(mesh) is ExplicitExerciseMesh:
        off = np.ascontiguousarray(mesh.schedules.offsets[:n + 1])
        times = np.ascontiguousarray(mesh.schedules.times[:int(off[-1] - off[0])])
        mesh = ExplicitExerciseMesh(schedules=RaggedExerciseTimes(offsets=off, times=times),
                                    origin_exercise=mesh.origin_exercise)
    counts = None if path_counts is None else RaggedPathCounts(
        training=np.ascontiguousarray(path_counts.training[:n]),
        evaluation=np.ascontiguousarray(path_counts.evaluation[:n]))
    per_group = [np.ascontiguousarray(a[:n]) for a in (spots, maturities, rates, vols)]
    return dict(group_offsets=go, initial_spots=per_group[0], strikes=np.ascontiguousarray(strikes[:rows]),
                maturities=per_group[1], rates=per_group[2], volatilities=per_group[3], dividends=div,
                mesh=mesh, path_counts=counts)


def _sample_groups(group_offsets, mesh, dividends, method, path_counts) -> int:
    groups = int(group_offsets.size - 1)
    strikes = np.diff(group_offsets.astype(np.int64)).astype(np.float64)
    rows = _group_rows(mesh, dividends, groups).astype(np.float64)
    if path_counts is not None:
        paths = path_counts.training.astype(np.float64) + path_counts.evaluation.astype(np.float64)
    else:
        paths = np.full(groups, float(method.training_paths + method.evaluation_paths))
    work = np.cumsum(rows * paths * (1.0 + strikes))
    target = min(work[-1] / SAMPLE_FRACTION, SAMPLE_WORK_CAP)
    return int(min(groups, max(1, np.searchsorted(work, target) + 1)))


def _bits(result) -> list:
    cols = [result.prices, result.standard_errors, result.early_exercise_paths, result.expiry_paths]
    if result.greeks:
        cols += [result.greeks[g] for g in sorted(result.greeks, key=lambda g: g.value)]
    return [np.asarray(c).tobytes() for c in cols]


def tune_request(executor, price: Callable, *, group_offsets, initial_spots, strikes, maturities,
                 rates, volatilities, dividends, mesh, method, greeks, path_counts) -> TuningRecord:
    """Choose and install the fastest invariant launch parameters for this request's shape.

    ``price(**sample)`` prices a request given as group columns (the batch module's untuned
    pricing function with the method and Greeks bound)."""
    key = shape_key(executor, group_offsets, mesh, dividends, method, greeks, path_counts)
    schedule, mib, fit = executor.tuning[2], executor.tuning[3], executor.tuning[4]
    rec = _CACHE.records.get(key)
    if rec is not None:
        executor.tune(**rec.choice, batch_schedule=schedule, batch_mib=mib, fit_schedule=fit)
        return rec
    n = _sample_groups(group_offsets, mesh, dividends, method, path_counts)
    sample = _slice(n, group_offsets, initial_spots, strikes, maturities, rates, volatilities,
                    dividends, mesh, path_counts)
    most_strikes = int(np.max(np.diff(group_offsets.astype(np.int64))))
    tiles = [t for t in (GREEK_TILES if greeks else PLAIN_TILES) if t <= most_strikes]
    from .simulation import HipPrecision

    fused = executor.precision is not HipPrecision.F64
    max_rows = int(np.max(_group_rows(mesh, dividends, int(group_offsets.size - 1))))
    encodings = ["value_columns"] + (["stopping_rows"] if fused and max_rows <= 255 else [])
    rec = TuningRecord(key=key, choice={}, sample_groups=n)
    reference = None

    def run(tile, cash):
        nonlocal reference
        executor.tune(eval_strike_tile=tile, fit_cashflows=cash, batch_schedule=schedule, batch_mib=mib,
                      fit_schedule=fit)
        out = price(**sample)   # warm-up: code objects, buffers
        bits = _bits(out)
        if reference is None:
            reference = bits
        elif bits != reference:
            raise ExecutionError(f"launch parameters (tile {tile}, {cash}) changed the results of a "
                                 "request whose results must not depend on them")
        best = float("inf")
        for _ in range(REPEATS):
            t0 = time.perf_counter()
            price(**sample)
            best = min(best, time.perf_counter() - t0)
        rec.timings[(tile, cash)] = best
        return best

    for tile in tiles:
        run(tile, "value_columns")
    best_tile = min((t for t in tiles), key=lambda t: rec.timings[(t, "value_columns")])
    for cash in encodings[1:]:
        run(best_tile, cash)
    tile, cash = min(rec.timings, key=rec.timings.get)
    rec.choice = {"eval_strike_tile": tile, "fit_cashflows": cash}
    _CACHE.records[key] = rec
    _CACHE.tuned += 1
    executor.tune(**rec.choice, batch_schedule=schedule, batch_mib=mib, fit_schedule=fit)
    return rec


__all__ = ["GREEK_TILES", "PLAIN_TILES", "TuningCache", "TuningRecord", "cache", "shape_key",
           "tune_request"]
