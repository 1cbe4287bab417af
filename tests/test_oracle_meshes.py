"""Pins of the oracle's event-grid builders (uniform, cash-dividend-date and explicit exercise
meshes) against the reference's own known answers and bitwise mesh invariants.

Sources:
* exact_union_times known answers, crates/finoptions-lsm/src/batch/pricing.rs:1318-1336;
* explicit mesh == uniform mesh with an AT_OR_BEFORE terminal cash dividend, bitwise, with three
  regression nodes (pricing.rs:1338-1390, tests/test_batch_lsm.py:523-584), and the sparse
  terminal-predecessor mesh with one regression node (tests/test_batch_lsm.py:565-584);
* cash-dividend-date mesh == UniformExerciseMesh(2) for one dividend at t = 0.5, bitwise
  (tests/test_batch_lsm.py:488-520);
* explicit [0.5], [0.375] == uniform(2) for maturities 1.0 and 0.75 (tests/test_batch_lsm.py:627-688).

A hand-built explicit union (exercise and cash times merged, equal times once, maturity last)
pins the merge itself: a dropped equal-time advance duplicates a node, and a non-strict maturity
test duplicates the last one.
"""

from __future__ import annotations

import numpy as np
import pytest

COLUMNS = ("prices", "sample_standard_deviations", "standard_errors", "origin_decisions_fitted",
           "origin_exercised", "regression_nodes", "no_training_support_nodes",
           "insufficient_exercise_support_nodes", "minimum_in_the_money_training_paths",
           "maximum_regression_condition_numbers", "early_exercise_paths", "expiry_paths")


def _request(oracle, *, spots=(100.0,), strikes=(100.0,), maturities=(1.0,), rates=(0.03,),
             vols=(0.25,), yields=None, cash=None, paths=2048, train_domain=1401,
             eval_domain=91401, **kw):
    n = len(spots)
    cols = dict(group_offsets=np.arange(n + 1, dtype=np.uint32), initial_spots=np.array(spots),
                strikes=np.array(strikes), maturities=np.array(maturities),
                rates=np.array(rates), volatilities=np.array(vols))
    if cash is not None:
        offsets, times, amounts = cash
        cols.update(cash_offsets=np.array(offsets, dtype=np.int64), cash_times=np.array(times),
                    cash_amounts=np.array(amounts))
    else:
        cols["dividend_yields"] = np.array(yields if yields is not None else [0.0] * n)
    # the reference tests' method(): training stream (7, d), evaluation stream (11, d')
    return oracle.GroupedRequest(training_paths=paths, evaluation_paths=paths, training_seed=7,
                                 evaluation_seed=11, training_domain_start=train_domain,
                                 evaluation_domain_start=eval_domain, **cols, **kw)


def _assert_equal(a: dict, b: dict) -> None:
    for name in COLUMNS:
        assert np.array_equal(np.asarray(a[name]), np.asarray(b[name])), name
        if np.asarray(a[name]).dtype == np.float64:
            assert np.array_equal(np.asarray(a[name]).view(np.uint64),
                                  np.asarray(b[name]).view(np.uint64)), name


def test_uniform_union_known_answers(oracle):
    """exact_union_times(0, 1.0, 4, [0.1, 0.5, 0.9]) and the grid point just below 0.5."""
    req = _request(oracle, cash=([0, 3], [0.1, 0.5, 0.9], [1.0, 1.0, 1.0]), mesh_kind=0,
                   base_intervals=4)
    grid = oracle.group_grid(req, 0)
    assert grid["times"].tolist() == [0.0, 0.1, 0.25, 0.5, 0.75, 0.9, 1.0]
    assert grid["dividend_nodes"].tolist() == [1, 3, 5]
    assert grid["exercise_nodes"].tolist() == [1, 2, 3, 4, 5]
    below_half = float(np.nextafter(0.5, 0.0))
    req = _request(oracle, cash=([0, 1], [below_half], [1.0]), mesh_kind=0, base_intervals=2)
    grid = oracle.group_grid(req, 0)
    assert grid["times"].tolist() == [0.0, below_half, 0.5, 1.0]
    assert grid["times"].view(np.uint64)[1] == np.float64(0.5).view(np.uint64) - 1


def test_uniform_union_at_the_largest_maturity_stays_finite(oracle):
    req = _request(oracle, maturities=(np.finfo(np.float64).max,), mesh_kind=0, base_intervals=4)
    times = oracle.group_grid(req, 0)["times"]
    assert times[-1] == np.finfo(np.float64).max
    assert np.all(np.isfinite(times)) and times.size == 5


def test_explicit_union_hand_built_grid(oracle):
    cash = ([0, 3], [0.1, 0.5, 0.9], [1.0, 2.0, 3.0])
    ex = dict(mesh_kind=2, exercise_offsets=np.array([0, 3], dtype=np.int64),
              exercise_times=np.array([0.2, 0.5, 0.7]))
    grid = oracle.group_grid(_request(oracle, cash=cash, origin_exercise=0, **ex), 0)
    assert grid["times"].tolist() == [0.0, 0.1, 0.2, 0.5, 0.7, 0.9, 1.0]
    assert grid["exercise_nodes"].tolist() == [2, 3, 4]
    assert grid["dividend_nodes"].tolist() == [1, 3, 5]
    assert grid["dividend_amounts"].tolist() == [1.0, 2.0, 3.0]
    grid = oracle.group_grid(_request(oracle, cash=cash, origin_exercise=1, **ex), 0)
    assert grid["exercise_nodes"].tolist() == [0, 2, 3, 4]
    # a terminal dividend (the AT_OR_BEFORE boundary) lands on the maturity node, added once
    terminal = ([0, 4], [0.1, 0.5, 0.9, 1.0], [1.0, 2.0, 3.0, 4.0])
    grid = oracle.group_grid(_request(oracle, cash=terminal, origin_exercise=0, **ex), 0)
    assert grid["times"].tolist() == [0.0, 0.1, 0.2, 0.5, 0.7, 0.9, 1.0]
    assert grid["dividend_nodes"].tolist() == [1, 3, 5, 6]
    # no cash: the exercise times and maturity alone
    grid = oracle.group_grid(_request(oracle, **ex), 0)
    assert grid["times"].tolist() == [0.0, 0.2, 0.5, 0.7, 1.0]
    assert grid["exercise_nodes"].tolist() == [1, 2, 3]


def test_cash_dividend_date_grid(oracle):
    """Base mesh of one interval united with the cash dates; every interior node is an exercise
    node and the origin never is."""
    req = _request(oracle, cash=([0, 2], [0.3, 0.6], [1.0, 2.0]), mesh_kind=1, origin_exercise=1)
    grid = oracle.group_grid(req, 0)
    assert grid["times"].tolist() == [0.0, 0.3, 0.6, 1.0]
    assert grid["exercise_nodes"].tolist() == [1, 2]
    assert grid["dividend_nodes"].tolist() == [1, 2]


def test_cash_dividend_date_mesh_equals_uniform_two_interval_mesh(oracle):
    """tests/test_batch_lsm.py:488-520: one dividend 2.0 at t = 0.5, liquidator, strike 95."""
    common = dict(strikes=(95.0,), cash=([0, 1], [0.5], [2.0]), cash_policy=0,
                  train_domain=913, eval_domain=90913, paths=4096)
    special = oracle.price_grouped(_request(oracle, mesh_kind=1, **common))
    uniform = oracle.price_grouped(_request(oracle, mesh_kind=0, base_intervals=2,
                                            origin_exercise=0, **common))
    _assert_equal(special, uniform)
    assert special["regression_nodes"][0] == 1


def test_explicit_mesh_equals_uniform_mesh_with_terminal_cash(oracle):
    """pricing.rs:1338-1390: cash 4.0 at maturity (AT_OR_BEFORE), uniform(4) vs explicit
    [0.25, 0.5, 0.75]; three regression nodes."""
    terminal = ([0, 1], [1.0], [4.0])
    uniform = oracle.price_grouped(_request(oracle, cash=terminal, mesh_kind=0, base_intervals=4))
    exact = oracle.price_grouped(_request(
        oracle, cash=terminal, mesh_kind=2, exercise_offsets=np.array([0, 3], dtype=np.int64),
        exercise_times=np.array([0.25, 0.5, 0.75])))
    _assert_equal(exact, uniform)
    assert exact["regression_nodes"][0] == 3
    # the terminal cash is paid before the expiry payoff: the price falls below the no-cash one
    plain = oracle.price_grouped(_request(oracle, mesh_kind=0, base_intervals=4))
    assert exact["prices"][0] < plain["prices"][0]


def test_explicit_terminal_predecessor_is_sparse(oracle):
    """tests/test_batch_lsm.py:565-584: one exercise date just before the terminal dividend."""
    req = _request(oracle, cash=([0, 1], [1.0], [4.0]), mesh_kind=2,
                   exercise_offsets=np.array([0, 1], dtype=np.int64),
                   exercise_times=np.array([255.0 / 256.0]))
    grid = oracle.group_grid(req, 0)
    assert grid["times"].tolist() == [0.0, 255.0 / 256.0, 1.0]
    assert grid["dividend_nodes"].tolist() == [2]
    serial = oracle.price_grouped(req, threads=1)
    parallel = oracle.price_grouped(req, threads=2)
    _assert_equal(serial, parallel)
    assert serial["regression_nodes"][0] == 1


def test_explicit_mesh_equals_uniform_for_rows_and_groups(oracle):
    """tests/test_batch_lsm.py:627-688: explicit [0.5] and [0.375] equal uniform(2) for maturities
    1.0 and 0.75, one row per group."""
    common = dict(spots=(100.0, 120.0), strikes=(100.0, 110.0), maturities=(1.0, 0.75),
                  rates=(0.05, 0.03), vols=(0.25, 0.30), yields=(0.02, 0.01),
                  train_domain=1601, eval_domain=91601)
    explicit = oracle.price_grouped(_request(
        oracle, mesh_kind=2, exercise_offsets=np.array([0, 1, 2], dtype=np.int64),
        exercise_times=np.array([0.5, 0.375]), **common), threads=2)
    uniform = oracle.price_grouped(_request(oracle, mesh_kind=0, base_intervals=2, **common),
                                   threads=2)
    _assert_equal(explicit, uniform)


@pytest.mark.parametrize("policy", [0, 1])
def test_cash_date_and_uniform_meshes_agree_for_both_policies(oracle, policy):
    """Two dividends on the union grid [0, 0.3, 0.6, 1]: the cash-date mesh equals an explicit
    mesh declaring the same dates, for the liquidator and the survivor policy; the two policies
    differ once a dividend exceeds some spots."""
    cash = ([0, 2], [0.3, 0.6], [60.0, 10.0])
    common = dict(cash=cash, strikes=(20.0,), vols=(0.6,), paths=4096)
    special = oracle.price_grouped(_request(oracle, mesh_kind=1, cash_policy=policy, **common))
    exact = oracle.price_grouped(_request(
        oracle, mesh_kind=2, exercise_offsets=np.array([0, 2], dtype=np.int64),
        exercise_times=np.array([0.3, 0.6]), cash_policy=policy, **common))
    _assert_equal(special, exact)
    other = oracle.price_grouped(_request(oracle, mesh_kind=1, cash_policy=1 - policy, **common))
    assert other["prices"][0] != special["prices"][0]
