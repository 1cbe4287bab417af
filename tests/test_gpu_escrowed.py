"""Escrowed cash dividends on the GPU (crates/finoptions-sim/src/path.rs:618-944,
crates/finoptions-lsm/src/batch/input.rs:798-860, pricing.rs:840-860; ADR 0011).

Bars as for the spot models: plain FP64 prices, standard errors and stopping counts within 1e-10 of
the oracle, the fused method within 0.01 SE, pathwise Greeks within 1e-9 of the oracle's (whose
chain rule through F0 and sigma_F is pinned by finite differences in test_oracle.py). Plus the
reference's own escrowed checks: spot matching is exactly the transformed prepaid-forward
volatility, a row batch equals the grouped batch, and a nonpositive F0 fails before simulation."""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method, oracle_grouped, rel_err

pytestmark = pytest.mark.gpu

PRICE_TOL = 1e-10


def _cols(convention, strikes=16):
    cols = c2_columns(strikes=strikes)
    del cols["yields"]
    cols.update(cash_offsets=np.array([0, 3], dtype=np.int64),
                cash_times=np.array([0.25, 0.5, 0.75]), cash_amounts=np.array([1.5, 1.5, 2.0]),
                escrowed=convention)
    return cols


def _pair(hip, oracle, cols, *, base_intervals, paths, seed, origin=False, greeks=()):
    import finoptions_hip as fh

    mesh = fh.UniformExerciseMesh(
        base_intervals=base_intervals,
        origin_exercise=fh.OriginExercise.INCLUDED if origin else fh.OriginExercise.EXCLUDED)
    method = hip_method(paths, paths, seed, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    res = fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=hip,
                                          method=method, greeks=greeks)
    req = oracle_grouped(cols, base_intervals=base_intervals, training_paths=paths,
                         evaluation_paths=paths, seed=seed, insufficient_support=1,
                         origin=1 if origin else 0)
    return res, req


@pytest.mark.parametrize("convention,origin", [(0, False), (1, False), (1, True)])
def test_escrowed_fp64_matches_the_oracle(hip_f64, oracle, convention, origin):
    cols = _cols(convention)
    res, req = _pair(hip_f64, oracle, cols, base_intervals=12, paths=16_384, seed=3, origin=origin)
    ref = oracle.price_grouped(req)
    err = rel_err(res.prices, ref["prices"])
    print(f"escrowed convention {convention} origin {origin}: max price rel err {err:.2e}")
    assert err < PRICE_TOL
    assert rel_err(res.standard_errors, ref["standard_errors"]) < PRICE_TOL
    for name in ("early_exercise_paths", "expiry_paths", "regression_nodes",
                 "minimum_in_the_money_training_paths", "origin_decisions_fitted",
                 "origin_exercised"):
        assert np.array_equal(getattr(res, name), ref[name]), name
    # a distinct model: the exact cash-jump request prices differently
    jump, jreq = _pair(hip_f64, oracle, {**cols, "escrowed": None}, base_intervals=12,
                       paths=16_384, seed=3, origin=origin)
    assert not np.array_equal(jump.prices, res.prices)


def test_escrowed_fused_method_matches_the_oracle(hip_fused, oracle):
    cols = _cols(1, strikes=32)
    res, req = _pair(hip_fused, oracle, cols, base_intervals=24, paths=32_768, seed=5)
    ref = oracle.price_grouped(req)
    bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
    print(f"escrowed fused vs oracle: max |bias|/SE {bias.max():.2e}")
    assert bias.max() < 0.01
    assert np.array_equal(res.expiry_paths, ref["expiry_paths"])


@pytest.mark.parametrize("convention", [0, 1])
def test_escrowed_pathwise_greeks_match_the_oracle(hip_f64, oracle, convention):
    import finoptions_hip as fh

    cols = _cols(convention, strikes=8)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)
    res, req = _pair(hip_f64, oracle, cols, base_intervals=8, paths=16_384, seed=9, greeks=greeks)
    ref = oracle.grouped_pathwise_greeks(req, 0)
    for g in greeks:
        err = rel_err(res.greeks[g], ref[g.value])
        print(f"escrowed {convention} {g.value}: max rel err vs oracle pathwise {err:.2e}")
        assert err < 1e-9


def _row_inputs(fh, volatility, convention, amount=10.0):
    return dict(
        initial_spots=np.array([100.0]), strikes=np.array([100.0]), maturities=np.array([1.0]),
        rates=np.array([0.0]), volatilities=np.array([volatility]),
        dividends=fh.EscrowedCashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=np.array([0, 1], dtype=np.int64),
                                             times=np.array([0.5]), amounts=np.array([amount])),
            volatility_convention=convention,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE),
        mesh=fh.UniformExerciseMesh(base_intervals=4, origin_exercise=fh.OriginExercise.EXCLUDED),
        method=hip_method(8192, 8192, 12, support=fh.InsufficientExerciseSupportPolicy.CONTINUE))


def test_spot_matching_is_exactly_the_transformed_prepaid_forward_volatility(hip_f64):
    """tests/test_batch_lsm.py:742-790 of the reference: sigma_F = sigma_S S0 / F0."""
    import finoptions_hip as fh

    matched = fh.price_lsm_call_batch(
        executor=hip_f64, **_row_inputs(fh, 0.20, fh.EscrowedVolatilityConvention.SPOT_MATCHED_AT_ORIGIN))
    direct = fh.price_lsm_call_batch(
        executor=hip_f64, **_row_inputs(fh, 0.20 * 100.0 / 90.0,
                                        fh.EscrowedVolatilityConvention.PREPAID_FORWARD))
    assert np.array_equal(matched.prices, direct.prices)
    assert np.array_equal(matched.standard_errors, direct.standard_errors)
    assert np.array_equal(matched.early_exercise_paths, direct.early_exercise_paths)


def test_nonpositive_prepaid_forward_fails_before_simulation(hip_f64):
    import finoptions_hip as fh

    with pytest.raises(fh.InputError, match=r"prepaid forward must be strictly positive.*row 0"):
        fh.price_lsm_call_batch(
            executor=hip_f64,
            **_row_inputs(fh, 0.20, fh.EscrowedVolatilityConvention.PREPAID_FORWARD, amount=100.0))


def test_escrowed_row_batch_equals_the_grouped_batch(hip_f64):
    """tests/test_batch_lsm.py:168-245 of the reference: one-strike groups are the row batch."""
    import finoptions_hip as fh

    schedules = fh.RaggedCashDividends(offsets=np.array([0, 2, 3], dtype=np.int64),
                                       times=np.array([0.2, 0.6, 0.4]),
                                       amounts=np.array([1.0, 1.0, 2.0]))
    dividends = fh.EscrowedCashDividendBatch(
        schedules=schedules,
        volatility_convention=fh.EscrowedVolatilityConvention.SPOT_MATCHED_AT_ORIGIN,
        maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    cols = dict(initial_spots=np.array([100.0, 80.0]), strikes=np.array([95.0, 85.0]),
                maturities=np.array([1.0, 0.75]), rates=np.array([0.03, 0.01]),
                volatilities=np.array([0.25, 0.35]))
    mesh = fh.UniformExerciseMesh(base_intervals=6, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(8192, 8192, 21, support=fh.InsufficientExerciseSupportPolicy.CONTINUE,
                        eval_domain=3)
    row = fh.price_lsm_call_batch(**cols, dividends=dividends, mesh=mesh, executor=hip_f64,
                                  method=method)
    grouped = fh.price_lsm_call_grouped_batch(
        groups=fh.LsmStrikeGroups(group_offsets=np.array([0, 1, 2], dtype=np.uint32),
                                  dividends=dividends, **cols),
        mesh=mesh, executor=hip_f64, method=method)
    assert np.array_equal(row.prices, grouped.prices)
    assert np.array_equal(row.standard_errors, grouped.standard_errors)
    assert np.array_equal(row.expiry_paths, grouped.expiry_paths)
    assert math.isfinite(float(row.prices[1]))
