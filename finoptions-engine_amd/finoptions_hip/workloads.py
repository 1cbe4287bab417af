"""Synthetic request columns for the engine's reference workloads.

Each builder returns plain NumPy columns (the exact layout the pricing functions borrow) so that
benchmarks, tests and the CPU checker share one definition of a workload.
"""

from __future__ import annotations

import numpy as np


def ragged_batch(contracts: int = 10_000, *, seed: int = 2024, min_exercise: int = 4,
                 max_exercise: int = 252, min_paths: int = 16_384, max_paths: int = 262_144,
                 path_scale: float = 1.0) -> dict:
    """Row-addressed ragged batch of Bermudan calls with cash dividends.

    Per contract (numpy ``default_rng(seed)``): n ~ U{min_exercise..max_exercise} equal intervals
    over T with exercise at T*i/n, i = 1..n-1 (plus expiry); 0-4 cash dividends at distinct
    randomly chosen exercise times with amounts U[0.1, 2.0]; S0 = exp(N(ln 100, 0.3));
    K = S0 * U[0.7, 1.3]; sigma ~ U[0.1, 0.6]; T ~ U[0.1, 3.0]; r ~ U[0, 0.05]; training =
    evaluation paths log-uniform in [min_paths, max_paths] (times ``path_scale``).
    Columns use Arrow-style int64 offsets for the exercise and cash schedules.
    """
    rng = np.random.default_rng(seed)
    n_ex = rng.integers(min_exercise, max_exercise + 1, size=contracts)
    spots = np.exp(rng.normal(np.log(100.0), 0.3, size=contracts))
    strikes = spots * rng.uniform(0.7, 1.3, size=contracts)
    vols = rng.uniform(0.1, 0.6, size=contracts)
    maturities = rng.uniform(0.1, 3.0, size=contracts)
    rates = rng.uniform(0.0, 0.05, size=contracts)
    log_paths = rng.uniform(np.log(min_paths), np.log(max_paths), size=contracts)
    paths = np.maximum(np.rint(np.exp(log_paths) * path_scale), 16).astype(np.uint32)
    ex_offsets = [0]
    ex_times: list[float] = []
    cash_offsets = [0]
    cash_times: list[float] = []
    cash_amounts: list[float] = []
    for c in range(contracts):
        n = int(n_ex[c])
        T = float(maturities[c])
        times = [T * (i / n) for i in range(1, n)]
        ex_times += times
        ex_offsets.append(len(ex_times))
        k = int(rng.integers(0, 5))
        k = min(k, len(times))
        chosen = np.sort(rng.choice(len(times), size=k, replace=False)) if k else np.array([], int)
        cash_times += [times[i] for i in chosen]
        cash_amounts += list(rng.uniform(0.1, 2.0, size=k))
        cash_offsets.append(len(cash_times))
    return {
        "initial_spots": spots, "strikes": strikes, "maturities": maturities, "rates": rates,
        "volatilities": vols,
        "exercise_offsets": np.array(ex_offsets, dtype=np.int64),
        "exercise_times": np.array(ex_times, dtype=np.float64),
        "cash_offsets": np.array(cash_offsets, dtype=np.int64),
        "cash_times": np.array(cash_times, dtype=np.float64),
        "cash_amounts": np.array(cash_amounts, dtype=np.float64),
        "paths": paths,
    }


def large_grouped_frame(groups: int = 4096, strikes_per_group: int = 256, *, seed: int = 5,
                        intervals: int = 12) -> dict:
    """Group-addressed frame: monthly mesh over T = 1 with quarterly cash dividends.

    Per group: S0 ~ U[20, 500], sigma ~ U[0.15, 0.5], r ~ U[0, 0.05], cash amounts ~ U[0.2, 1.0]
    at t = 0.25, 0.5, 0.75 (which coincide with mesh nodes 3, 6, 9); strikes S0 *
    linspace(0.8, 1.2, strikes_per_group); uint32 group offsets.
    """
    rng = np.random.default_rng(seed)
    spots = rng.uniform(20.0, 500.0, size=groups)
    vols = rng.uniform(0.15, 0.5, size=groups)
    rates = rng.uniform(0.0, 0.05, size=groups)
    amounts = rng.uniform(0.2, 1.0, size=(groups, 3))
    grid = np.linspace(0.8, 1.2, strikes_per_group)
    strikes = (spots[:, None] * grid[None, :]).reshape(-1)
    quarter = [1.0 * (3 / intervals), 1.0 * (6 / intervals), 1.0 * (9 / intervals)]
    return {
        "group_offsets": (np.arange(groups + 1, dtype=np.uint64) * strikes_per_group).astype(np.uint32),
        "initial_spots": spots, "strikes": strikes, "maturities": np.ones(groups),
        "rates": rates, "volatilities": vols,
        "cash_offsets": (np.arange(groups + 1, dtype=np.int64) * 3),
        "cash_times": np.tile(np.array(quarter), groups),
        "cash_amounts": amounts.reshape(-1),
        "intervals": intervals,
    }


def ragged_path_steps(w: dict) -> np.ndarray:
    """Simulated path-steps of every contract of a ragged batch: intervals x (training +
    evaluation paths), the exercise times plus expiry being the grid (the cash dates lie on
    exercise times)."""
    intervals = np.diff(w["exercise_offsets"]) + 1
    return intervals.astype(np.float64) * 2.0 * w["paths"].astype(np.float64)


def ragged_batch_request(w: dict, *, first: int = 0, count: int | None = None,
                         spot_scale: float = 1.0, vol_shift: float = 0.0, seed: int = 2024,
                         support=None) -> dict:
    """Keyword arguments of ``price_lsm_call_batch`` for contracts [first, first + count) of a
    ragged batch, with the stream domains those rows have in the whole batch (training
    1 + row, evaluation 2 + rows + row), liquidator cash dividends strictly before maturity, the
    explicit exercise mesh and per-contract path counts. ``spot_scale`` and ``vol_shift`` bump
    every contract's spot (multiplicatively) and volatility (additively, relative to itself)."""
    from . import (CashDividendBatch, DividendPolicy, ExplicitExerciseMesh,
                   InsufficientExerciseSupportPolicy, Lsm, MaturityBoundary, OriginExercise,
                   RaggedCashDividends, RaggedExerciseTimes, RaggedPathCounts, RandomStream)

    rows = len(w["initial_spots"])
    count = rows - first if count is None else count
    sl = slice(first, first + count)
    eo = w["exercise_offsets"][first:first + count + 1]
    co = w["cash_offsets"][first:first + count + 1]
    paths = np.ascontiguousarray(w["paths"][sl])
    method = Lsm(training_paths=int(paths.max()), evaluation_paths=int(paths.max()),
                 training_stream=RandomStream(seed=seed, domain=1 + first),
                 evaluation_stream=RandomStream(seed=seed, domain=2 + rows + first),
                 insufficient_exercise_support=support or InsufficientExerciseSupportPolicy.CONTINUE)
    vols = np.ascontiguousarray(w["volatilities"][sl])
    return dict(
        initial_spots=np.ascontiguousarray(w["initial_spots"][sl] * spot_scale),
        strikes=np.ascontiguousarray(w["strikes"][sl]),
        maturities=np.ascontiguousarray(w["maturities"][sl]),
        rates=np.ascontiguousarray(w["rates"][sl]),
        volatilities=np.ascontiguousarray(vols + vol_shift * vols) if vol_shift else vols,
        dividends=CashDividendBatch(
            schedules=RaggedCashDividends(offsets=np.ascontiguousarray(co),
                                          times=np.ascontiguousarray(w["cash_times"][co[0]:co[-1]]),
                                          amounts=np.ascontiguousarray(w["cash_amounts"][co[0]:co[-1]])),
            policy=DividendPolicy.LIQUIDATOR, maturity_boundary=MaturityBoundary.STRICTLY_BEFORE),
        mesh=ExplicitExerciseMesh(
            schedules=RaggedExerciseTimes(offsets=np.ascontiguousarray(eo),
                                          times=np.ascontiguousarray(w["exercise_times"][eo[0]:eo[-1]])),
            origin_exercise=OriginExercise.EXCLUDED),
        method=method,
        path_counts=RaggedPathCounts(training=paths, evaluation=paths))
