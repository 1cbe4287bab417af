"""The escrowed cash-dividend model in the oracle (crates/finoptions-sim/src/path.rs:618-944,
crates/finoptions-lsm/src/batch/input.rs:798-860, pricing.rs:840-860).

The observed stock is S = F + A(t): F, the risky prepaid forward, follows GBM with no dividends,
and A(t) is the prepaid value of the remaining cash. Pins, independent of the recursion under test:
the closed form of A, the identity "escrowed paths = dividend-free spot paths from F0 plus A", the
terminal-cash exclusion at expiry, the European closed form on F0, the exact equivalence of the
spot-matched convention with its transformed prepaid-forward volatility, and the early failure of
a nonpositive F0."""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import oracle_grouped


def _cols(cash_times, cash_amounts, *, strikes=(90.0, 100.0, 110.0), spot=100.0, rate=0.03,
          vol=0.25, maturity=1.0, escrowed=0):
    return {
        "group_offsets": np.array([0, len(strikes)], dtype=np.uint32),
        "initial_spots": np.array([spot]),
        "strikes": np.array(strikes, dtype=np.float64),
        "maturities": np.array([maturity]),
        "rates": np.array([rate]),
        "volatilities": np.array([vol]),
        "cash_offsets": np.array([0, len(cash_times)], dtype=np.int64),
        "cash_times": np.array(cash_times, dtype=np.float64),
        "cash_amounts": np.array(cash_amounts, dtype=np.float64),
        "escrowed": escrowed,
    }


def _origin_prepaid(times, amounts, rate):
    """input.rs:834-860: the schedule-order sum of amount * exp(-rate * time)."""
    total = 0.0
    for t, d in zip(times, amounts):
        total += d * math.exp(-rate * t)
    return total


def test_prepaid_cash_is_the_discounted_remaining_cash(oracle):
    times, amounts, rate = [0.25, 0.5, 1.0], [1.0, 2.0, 3.0], 0.04
    req = oracle_grouped(_cols(times, amounts, rate=rate), base_intervals=8, training_paths=8,
                         evaluation_paths=8, seed=1)
    grid = oracle.group_grid(req, 0)["times"]
    prepaid = oracle.escrowed_prepaid_cash(req, 0)
    assert prepaid.size == grid.size
    assert prepaid[0] == _origin_prepaid(times, amounts, rate)
    for n in range(1, grid.size):
        t = grid[n]
        expected = sum(d * math.exp(-rate * (tau - t)) for tau, d in zip(times, amounts) if tau >= t)
        assert prepaid[n] == pytest.approx(expected, rel=1e-14, abs=0.0), n
    # the terminal (maturity) dividend stays in every earlier node's value
    assert prepaid[-1] == 3.0


@pytest.mark.parametrize("convention", [0, 1])
def test_escrowed_paths_are_dividend_free_paths_from_f0_plus_prepaid_cash(oracle, convention):
    """With the cash dates also declared as exercise dates, the escrowed plan and a zero-yield
    spot plan started at F0 with volatility sigma_F share one grid and one stream, so every
    observation must equal the spot path plus A(node) bit for bit (and S0 at the origin)."""
    times, amounts, rate, spot, vol = [0.3, 0.6], [2.0, 3.0], 0.05, 100.0, 0.2
    ex = np.array([0.3, 0.45, 0.6, 0.8])
    cols = _cols(times, amounts, rate=rate, spot=spot, vol=vol, escrowed=convention)
    mesh = dict(mesh_kind=2, exercise_offsets=np.array([0, ex.size], dtype=np.int64),
                exercise_times=ex, origin_exercise=1)
    req = oracle_grouped(cols, base_intervals=1, training_paths=256, evaluation_paths=256, seed=3,
                         **mesh)
    a0 = _origin_prepaid(times, amounts, rate)
    f0 = spot - a0
    sigma_f = vol * spot / f0 if convention == 1 else vol
    spot_cols = {k: v for k, v in cols.items() if not k.startswith("cash") and k != "escrowed"}
    spot_cols.update(initial_spots=np.array([f0]), volatilities=np.array([sigma_f]),
                     yields=np.array([0.0]))
    spot_req = oracle_grouped(spot_cols, base_intervals=1, training_paths=256,
                              evaluation_paths=256, seed=3, **mesh)
    grid = oracle.group_grid(req, 0)
    assert np.array_equal(grid["times"], oracle.group_grid(spot_req, 0)["times"])
    prepaid = oracle.escrowed_prepaid_cash(req, 0)
    esc = oracle.grouped_paths(req, 0)
    ref = oracle.grouped_paths(spot_req, 0)
    obs_nodes = list(grid["exercise_nodes"]) + [grid["times"].size - 1]
    assert obs_nodes[0] == 0
    assert np.all(esc[0] == spot)
    for row, node in enumerate(obs_nodes[1:], start=1):
        add = 0.0 if node == grid["times"].size - 1 else prepaid[node]
        assert np.array_equal(esc[row], ref[row] + add), row
    # at a dividend date the observed stock includes that date's cash (pre-event observation)
    node_03 = int(np.searchsorted(grid["times"], 0.3))
    assert prepaid[node_03] == pytest.approx(2.0 + 3.0 * math.exp(-rate * 0.3), rel=1e-15)


def test_expiry_observation_excludes_paid_terminal_cash(oracle):
    """path.rs:1232-1242: a dividend at maturity is in A before expiry but not in the expiry
    observation, which is F_T itself."""
    cols = _cols([1.0], [10.0], rate=0.0, escrowed=0)
    req = oracle_grouped(cols, base_intervals=2, training_paths=64, evaluation_paths=64, seed=4)
    prepaid = oracle.escrowed_prepaid_cash(req, 0)
    assert prepaid.tolist() == [10.0, 10.0, 10.0]
    spot_cols = {k: v for k, v in cols.items() if not k.startswith("cash") and k != "escrowed"}
    spot_cols.update(initial_spots=np.array([90.0]), yields=np.array([0.0]))
    spot_req = oracle_grouped(spot_cols, base_intervals=2, training_paths=64, evaluation_paths=64,
                              seed=4)
    esc = oracle.grouped_paths(req, 0)
    ref = oracle.grouped_paths(spot_req, 0)
    assert np.array_equal(esc[-1], ref[-1])
    assert np.array_equal(esc[0], ref[0] + 10.0)


def _black_scholes_call(s, k, r, sigma, t):
    sd = sigma * math.sqrt(t)
    d1 = (math.log(s / k) + (r + 0.5 * sigma * sigma) * t) / sd
    d2 = d1 - sd
    n = lambda x: 0.5 * (1.0 + math.erf(x / math.sqrt(2.0)))
    return s * n(d1) - k * math.exp(-r * t) * n(d2)


def test_european_payoffs_converge_to_the_closed_form_on_the_prepaid_forward(oracle):
    """path.rs:1256-1300: with no exercise rows the price is the discounted terminal mean, and the
    terminal stock is F_T, so Black-Scholes with underlying F0 and volatility sigma_F holds."""
    rate, vol, spot, strike = 0.03, 0.25, 100.0, 100.0
    times, amounts = [0.4, 0.8], [2.0, 2.0]
    cols = _cols(times, amounts, strikes=(strike,), rate=rate, vol=vol, escrowed=0)
    req = oracle_grouped(cols, base_intervals=1, training_paths=64, evaluation_paths=262_144,
                         seed=0x0F317A115EEC2026, mesh_kind=2,
                         exercise_offsets=np.array([0, 0], dtype=np.int64),
                         exercise_times=np.zeros(0))
    res = oracle.price_grouped(req, threads=8)
    f0 = spot - _origin_prepaid(times, amounts, rate)
    expected = _black_scholes_call(f0, strike, rate, vol, 1.0)
    z = abs(res["prices"][0] - expected) / res["standard_errors"][0]
    assert z < 5.0, (res["prices"][0], expected, res["standard_errors"][0])
    # the same request on the exact cash-jump model is a different price (distinct model identity)
    jump = oracle.price_grouped(oracle_grouped({**cols, "escrowed": None}, base_intervals=1,
                                               training_paths=64, evaluation_paths=262_144,
                                               seed=0x0F317A115EEC2026, mesh_kind=2,
                                               exercise_offsets=np.array([0, 0], dtype=np.int64),
                                               exercise_times=np.zeros(0)), threads=8)
    assert jump["prices"][0] != res["prices"][0]


def test_spot_matched_volatility_is_exactly_the_transformed_prepaid_volatility(oracle):
    """ADR 0011 / tests/test_batch_lsm.py:742-790: spot matching is sigma_F = sigma_S S0 / F0, so a
    request that supplies that value under the prepaid-forward convention is identical."""
    times, amounts, rate, spot, vol = [0.5], [10.0], 0.0, 100.0, 0.2
    matched = oracle_grouped(_cols(times, amounts, strikes=(100.0,), rate=rate, spot=spot, vol=vol,
                                   escrowed=1),
                             base_intervals=4, training_paths=4096, evaluation_paths=4096, seed=8,
                             insufficient_support=1)
    f0 = spot - _origin_prepaid(times, amounts, rate)
    direct = oracle_grouped(_cols(times, amounts, strikes=(100.0,), rate=rate, spot=spot,
                                  vol=vol * spot / f0, escrowed=0),
                            base_intervals=4, training_paths=4096, evaluation_paths=4096, seed=8,
                            insufficient_support=1)
    a, b = oracle.price_grouped(matched), oracle.price_grouped(direct)
    for key in a:
        assert np.array_equal(a[key], b[key]), key
    assert f0 == 90.0 and vol * spot / f0 == pytest.approx(0.2 * 100.0 / 90.0, rel=0, abs=0)


def test_nonpositive_prepaid_forward_fails_before_simulation(oracle):
    req = oracle_grouped(_cols([0.5], [100.0], rate=0.0, escrowed=0), base_intervals=2,
                         training_paths=8, evaluation_paths=8, seed=1)
    with pytest.raises(oracle.OracleError) as info:
        oracle.price_grouped(req)
    assert info.value.code == 508
    assert info.value.phase == 0
