"""Full-size ragged-batch and large-frame requests, checked through properties that hold at any
size: every row's stopping counts partition its evaluation paths, every price is finite with a
positive standard error, a unit priced inside the big request is bitwise identical to the same
unit priced alone (per-unit arithmetic and reduction trees do not depend on batching), and
sampled units agree with the oracle. Throughput is printed for the record."""

from __future__ import annotations

import time

import numpy as np
import pytest

from conftest import oracle_grouped

pytestmark = pytest.mark.gpu


def _ragged_inputs(fh, w, rows, *, first=0, count=None):
    """Row batch over contracts [first, first + count) with the domains they have in the full
    request (training 1 + row, evaluation 2 + rows + row)."""
    count = rows - first if count is None else count
    sl = slice(first, first + count)
    eo = w["exercise_offsets"][first:first + count + 1]
    co = w["cash_offsets"][first:first + count + 1]
    dividends = fh.CashDividendBatch(
        schedules=fh.RaggedCashDividends(offsets=co, times=w["cash_times"][co[0]:co[-1]],
                                         amounts=w["cash_amounts"][co[0]:co[-1]]),
        policy=fh.DividendPolicy.LIQUIDATOR, maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    mesh = fh.ExplicitExerciseMesh(
        schedules=fh.RaggedExerciseTimes(offsets=eo, times=w["exercise_times"][eo[0]:eo[-1]]),
        origin_exercise=fh.OriginExercise.EXCLUDED)
    paths = np.ascontiguousarray(w["paths"][sl])
    method = fh.Lsm(training_paths=int(paths.max()), evaluation_paths=int(paths.max()),
                    training_stream=fh.RandomStream(seed=2024, domain=1 + first),
                    evaluation_stream=fh.RandomStream(seed=2024, domain=2 + rows + first),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return dict(initial_spots=np.ascontiguousarray(w["initial_spots"][sl]),
                strikes=np.ascontiguousarray(w["strikes"][sl]),
                maturities=np.ascontiguousarray(w["maturities"][sl]),
                rates=np.ascontiguousarray(w["rates"][sl]),
                volatilities=np.ascontiguousarray(w["volatilities"][sl]),
                dividends=dividends, mesh=mesh, method=method,
                path_counts=fh.RaggedPathCounts(training=paths, evaluation=paths))


def test_full_ragged_batch_of_10000_contracts(hip_fused, oracle):
    import finoptions_hip as fh
    from finoptions_hip.workloads import ragged_batch

    w = ragged_batch(10_000, seed=2024)
    rows = len(w["initial_spots"])
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    t0 = time.perf_counter()
    res = fh.price_lsm_call_batch(executor=hip_fused, greeks=greeks, **_ragged_inputs(fh, w, rows))
    seconds = time.perf_counter() - t0
    intervals = np.diff(w["exercise_offsets"]) + 1
    steps = float(np.sum(intervals * 2.0 * w["paths"]))
    print(f"ragged batch 10000 contracts: {seconds:.2f} s, {steps:.3e} path-steps, "
          f"{steps / seconds:.3e} path-steps/s (fused FP64, delta+vega)")
    assert np.array_equal(res.early_exercise_paths + res.expiry_paths, w["paths"])
    # a contract whose every payoff is zero has price 0 and standard error 0
    assert np.all(np.isfinite(res.prices)) and np.all(res.standard_errors >= 0)
    assert np.all((res.standard_errors > 0) | (res.prices == 0))
    print(f"contracts with no positive payoff: {int(np.sum(res.prices == 0))}")
    assert np.all(np.isfinite(res.greeks[fh.Greek.DELTA]))
    # batch invariance: contracts priced alone are bitwise identical to their rows
    for first in (0, 4_321, 9_999):
        alone = fh.price_lsm_call_batch(executor=hip_fused, greeks=greeks,
                                        **_ragged_inputs(fh, w, rows, first=first, count=1))
        assert alone.prices[0] == res.prices[first]
        assert alone.standard_errors[0] == res.standard_errors[first]
        assert alone.greeks[fh.Greek.DELTA][0] == res.greeks[fh.Greek.DELTA][first]
    # the oracle on the three contracts with the fewest paths
    pick = np.argsort(w["paths"])[:3]
    for row in pick:
        eo = w["exercise_offsets"][row:row + 2]
        co = w["cash_offsets"][row:row + 2]
        n = int(w["paths"][row])
        req = oracle.GroupedRequest(
            group_offsets=np.array([0, 1], dtype=np.uint32),
            initial_spots=w["initial_spots"][row:row + 1], strikes=w["strikes"][row:row + 1],
            maturities=w["maturities"][row:row + 1], rates=w["rates"][row:row + 1],
            volatilities=w["volatilities"][row:row + 1],
            cash_offsets=co, cash_times=w["cash_times"][co[0]:co[1]],
            cash_amounts=w["cash_amounts"][co[0]:co[1]], cash_policy=0,
            training_paths=n, evaluation_paths=n, training_seed=2024, evaluation_seed=2024,
            training_domain_start=1 + int(row), evaluation_domain_start=2 + rows + int(row),
            mesh_kind=2, exercise_offsets=eo, exercise_times=w["exercise_times"][eo[0]:eo[1]],
            origin_exercise=0, insufficient_support=1)
        ref = oracle.price_grouped(req)
        bias = abs(res.prices[row] - ref["prices"][0]) / ref["standard_errors"][0]
        print(f"contract {row}: {n} paths, {int(eo[1] - eo[0]) + 1} intervals, |bias|/SE {bias:.2e}")
        assert bias < 0.01
        assert res.expiry_paths[row] == ref["expiry_paths"][0]


def _frame_groups(fh, w, first, count):
    offs = w["group_offsets"][first:first + count + 1]
    co = w["cash_offsets"][first:first + count + 1]
    return fh.LsmStrikeGroups(
        group_offsets=np.ascontiguousarray(offs - offs[0]),
        initial_spots=np.ascontiguousarray(w["initial_spots"][first:first + count]),
        strikes=np.ascontiguousarray(w["strikes"][offs[0]:offs[-1]]),
        maturities=np.ascontiguousarray(w["maturities"][first:first + count]),
        rates=np.ascontiguousarray(w["rates"][first:first + count]),
        volatilities=np.ascontiguousarray(w["volatilities"][first:first + count]),
        dividends=fh.CashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=co, times=w["cash_times"][co[0]:co[-1]],
                                             amounts=w["cash_amounts"][co[0]:co[-1]]),
            policy=fh.DividendPolicy.LIQUIDATOR,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE))


def test_full_grouped_frame_4096_states_by_256_strikes(hip_fused, oracle):
    import finoptions_hip as fh
    from finoptions_hip.workloads import large_grouped_frame

    groups, per, paths = 4096, 256, 32_768
    w = large_grouped_frame(groups=groups, strikes_per_group=per)
    mesh = fh.UniformExerciseMesh(base_intervals=w["intervals"],
                                  origin_exercise=fh.OriginExercise.EXCLUDED)

    def method(first):
        return fh.Lsm(training_paths=paths, evaluation_paths=paths,
                      training_stream=fh.RandomStream(seed=5, domain=1 + first),
                      evaluation_stream=fh.RandomStream(seed=5, domain=1 + groups + first),
                      insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)

    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    t0 = time.perf_counter()
    res = fh.price_lsm_call_grouped_batch(groups=_frame_groups(fh, w, 0, groups), mesh=mesh,
                                          executor=hip_fused, method=method(0), greeks=greeks)
    seconds = time.perf_counter() - t0
    steps = float(groups) * w["intervals"] * 2 * paths
    print(f"grouped frame 4096 x 256: {seconds:.2f} s, {steps:.3e} path-steps, "
          f"{steps / seconds:.3e} path-steps/s, {groups * per / seconds:.3e} strike rows/s "
          f"(fused FP64, delta+vega)")
    assert np.all(res.early_exercise_paths + res.expiry_paths == paths)
    assert np.all(np.isfinite(res.prices)) and np.all(res.standard_errors > 0)
    assert np.all(np.isfinite(res.greeks[fh.Greek.VEGA]))
    for first in (0, 2_047, 4_095):
        alone = fh.price_lsm_call_grouped_batch(groups=_frame_groups(fh, w, first, 1), mesh=mesh,
                                                executor=hip_fused, method=method(first),
                                                greeks=greeks)
        rows = slice(first * per, (first + 1) * per)
        assert np.array_equal(alone.prices, res.prices[rows])
        assert np.array_equal(alone.greeks[fh.Greek.VEGA], res.greeks[fh.Greek.VEGA][rows])
    # one whole group against the oracle
    g = 2_047
    cols = {k: w[k] for k in ("initial_spots", "maturities", "rates", "volatilities")}
    cols = {k: v[g:g + 1] for k, v in cols.items()}
    cols["group_offsets"] = np.array([0, per], dtype=np.uint32)
    cols["strikes"] = w["strikes"][g * per:(g + 1) * per]
    cols["cash_offsets"] = np.array([0, 3], dtype=np.int64)
    cols["cash_times"] = w["cash_times"][3 * g:3 * g + 3]
    cols["cash_amounts"] = w["cash_amounts"][3 * g:3 * g + 3]
    req = oracle_grouped(cols, base_intervals=w["intervals"], training_paths=paths,
                         evaluation_paths=paths, seed=5, train_domain=1 + g,
                         eval_domain=1 + groups + g, insufficient_support=1)
    ref = oracle.price_grouped(req, threads=16)
    bias = np.abs(res.prices[g * per:(g + 1) * per] - ref["prices"]) / ref["standard_errors"]
    print(f"group {g} vs oracle: max |bias|/SE {bias.max():.2e}")
    assert bias.max() < 0.01


def _fixture_columns(name: str) -> dict:
    import json
    from pathlib import Path

    raw = json.loads((Path(__file__).parent / "golden" / name).read_text())
    out = {}
    for k, v in raw.items():
        if isinstance(v, list):
            out[k] = np.array([float.fromhex(x) for x in v]) if isinstance(v[0], str) else np.array(v)
        else:
            out[k] = v
    return out


def test_c2_full_size_matches_the_oracle_fixture(hip_f64, hip_fused):
    """The grouped-strike state at its full size (64 strikes x 50 intervals, 2^20 + 2^20 paths)
    against the oracle's columns committed by tests/golden/make_c2_oracle_fixture.py. The FP64
    method must agree to 1e-10 relative on every strike whose stopping counts agree exactly, and a
    strike whose counts differ (a decision flipped at a near tie by a last-bit difference of the
    device's exp) is reported with its shift in standard errors; the fused method stays within a
    hundredth of a standard error everywhere."""
    import finoptions_hip as fh

    ref = _fixture_columns("c2_full_oracle.json")
    paths = int(ref["paths"])
    cols_kw = dict(group_offsets=np.array([0, 64], dtype=np.uint32), initial_spots=np.array([100.0]),
                   strikes=np.linspace(80.0, 120.0, 64), maturities=np.array([1.0]),
                   rates=np.array([0.03]), volatilities=np.array([0.25]))
    groups = fh.LsmStrikeGroups(**cols_kw, dividends=fh.ContinuousDividendYieldBatch(
        annualized_decimals=np.array([0.02])))
    mesh = fh.UniformExerciseMesh(base_intervals=50, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = fh.Lsm(training_paths=paths, evaluation_paths=paths,
                    training_stream=fh.RandomStream(seed=11, domain=1),
                    evaluation_stream=fh.RandomStream(seed=11, domain=2),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    for hip, label in ((hip_f64, "F64"), (hip_fused, "fused")):
        res = fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=hip, method=method)
        same = ((res.early_exercise_paths == ref["early_exercise_paths"])
                & (res.expiry_paths == ref["expiry_paths"]))
        rel = np.abs(res.prices - ref["prices"]) / np.abs(ref["prices"])
        rel_se = np.abs(res.standard_errors - ref["standard_errors"]) / ref["standard_errors"]
        bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
        print(f"C2 full size {label}: strikes with equal stopping counts {int(same.sum())}/64; "
              f"max price rel err there {rel[same].max() if same.any() else 0:.2e}; "
              f"max |bias|/SE {bias.max():.2e}; flipped strikes {np.flatnonzero(~same).tolist()}")
        assert bias.max() < 0.01
        assert np.array_equal(res.regression_nodes, ref["regression_nodes"])
        if label == "F64":
            assert rel[same].max() < 1e-10
            assert rel_se[same].max() < 1e-10
            assert int((~same).sum()) <= 2
