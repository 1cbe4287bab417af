"""Tie counts on the GPU against the oracle's (tests/test_oracle_ties.py pins the oracle's with an
exact-tie case and an independent recount).

A group with volatility 1e-9 and degree-0 rules puts every in-the-money decision within 2^-20 of
its continuation: with that window both methods must count every such decision, as the oracle
does. With a window of 2^-8 over a C2 slice the plain FP64 method, whose continuations are the
reference's, must count exactly the oracle's near ties, and the fused method must agree up to the
decisions whose distance to the window edge lies inside the last-bit difference of the two
continuations. The window is a reporting parameter: prices never depend on it. (The default,
window 0, counts exact ties, which continuous prices make vanishingly rare; the API rejects the
zero volatility that would force them.)"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method, oracle_grouped

pytestmark = pytest.mark.gpu

TIES = ("near_tie_training_decisions", "near_tie_evaluation_decisions")


def _flat_columns():
    return dict(group_offsets=np.array([0, 4], dtype=np.uint32), initial_spots=np.array([110.0]),
                strikes=np.array([90.0, 100.0, 105.0, 120.0]), maturities=np.array([1.0]),
                rates=np.array([0.0]), volatilities=np.array([1e-9]), yields=np.array([0.0]))


def _price(executor, cols, *, intervals, paths, seed, degree=4, greeks=()):
    import finoptions_hip as fh

    method = hip_method(paths, paths, seed, support=fh.InsufficientExerciseSupportPolicy.CONTINUE,
                        degree=degree)
    mesh = fh.UniformExerciseMesh(base_intervals=intervals, origin_exercise=fh.OriginExercise.EXCLUDED)
    return fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


@pytest.mark.parametrize("which", ["hip_f64", "hip_fused"])
def test_every_decision_within_the_window_is_counted(request, oracle, which):
    hip = request.getfixturevalue(which)
    cols = _flat_columns()
    paths, window = 8192, 2.0 ** -20
    req = oracle_grouped(cols, base_intervals=2, training_paths=paths, evaluation_paths=paths, seed=5,
                         insufficient_support=1, degree=0, tie_window=window)
    ref = oracle.price_grouped(req)
    hip.set_near_tie_window(window)
    try:
        res = _price(hip, cols, intervals=2, paths=paths, seed=5, degree=0)
    finally:
        hip.set_near_tie_window()
    print(f"{which}: ties {res.near_tie_training_decisions}, {res.near_tie_evaluation_decisions}")
    for name in TIES:
        assert np.array_equal(getattr(res, name), ref[name]), name
    itm = cols["strikes"] < 110.0
    assert np.all(res.near_tie_training_decisions[itm] == paths)
    assert np.all(res.near_tie_evaluation_decisions[itm] == paths)
    assert np.all(res.near_tie_training_decisions[~itm] == 0)
    assert np.all(res.near_tie_evaluation_decisions[~itm] == 0)
    # the default counts exact ties only
    exact = _price(hip, cols, intervals=2, paths=paths, seed=5, degree=0)
    assert np.array_equal(exact.prices.view(np.uint64), res.prices.view(np.uint64))
    assert np.all(exact.near_tie_evaluation_decisions <= res.near_tie_evaluation_decisions)


def test_wide_window_counts_match_the_oracle(hip_f64, hip_fused, oracle):
    cols = c2_columns(strikes=16)
    window = 2.0 ** -8
    req = oracle_grouped(cols, base_intervals=10, training_paths=16_384, evaluation_paths=16_384,
                         seed=13, insufficient_support=1, tie_window=window)
    ref = oracle.price_grouped(req)
    default = _price(hip_fused, cols, intervals=10, paths=16_384, seed=13)
    for hip in (hip_f64, hip_fused):
        hip.set_near_tie_window(window)
    try:
        f64 = _price(hip_f64, cols, intervals=10, paths=16_384, seed=13)
        fused = _price(hip_fused, cols, intervals=10, paths=16_384, seed=13)
    finally:
        for hip in (hip_f64, hip_fused):
            hip.set_near_tie_window()
    for name in TIES:
        diff = np.abs(fused.__getattribute__(name).astype(np.int64) - ref[name].astype(np.int64))
        print(f"{name}: oracle total {int(ref[name].sum())}, F64 GPU {int(getattr(f64, name).sum())}, "
              f"fused GPU {int(getattr(fused, name).sum())} (max row difference {int(diff.max())})")
        assert ref[name].sum() > 0
        assert np.array_equal(getattr(f64, name), ref[name]), name
        assert diff.max() <= 2, name
    # the window is a reporting parameter: every other column is unchanged
    assert np.array_equal(fused.prices.view(np.uint64), default.prices.view(np.uint64))
    assert np.array_equal(fused.early_exercise_paths, default.early_exercise_paths)
    # the default counts exact ties, as the oracle does at window 0
    exact = oracle.price_grouped(oracle_grouped(cols, base_intervals=10, training_paths=16_384,
                                                evaluation_paths=16_384, seed=13,
                                                insufficient_support=1, tie_window=0.0))
    for name in TIES:
        print(f"exact {name}: oracle {int(exact[name].sum())}, fused GPU {int(getattr(default, name).sum())}")
        assert np.array_equal(getattr(default, name), exact[name]), name
