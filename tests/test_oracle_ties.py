"""Near-tie counts of the oracle (this project's decision-flip accounting; the reference has no
such diagnostic, so it is pinned here by a closed-form case and by an independent recount).

A rule application is a near tie when |intrinsic - continuation| <= window |continuation|. With
window 0 the count is the number of exact ties, which are exactly the decisions the tie rule
settles: a zero-volatility plan makes every decision an exact tie, so the two tie rules must then
stop every path at opposite rows while counting every decision. With a wide window the counts must
equal a recount in numpy that recomputes every continuation from the exported rules with its own
three-term recurrence and replays the evaluation's stopping."""

from __future__ import annotations

import numpy as np
import pytest


def test_exact_ties_are_counted_and_settled_by_the_tie_rule(oracle):
    plan = oracle.Plan(times=[0.0, 0.5, 1.0], log_discounts=[0.0, 0.0, 0.0],
                       cum_yields=[0.0, 0.0, 0.0], volatility=0.0, exercise_nodes=[1],
                       expiry_node=2)
    n = 16
    z = np.zeros((2, n))
    training = oracle.simulate_normals(plan, 110.0, z)
    evaluation = oracle.simulate_normals(plan, 110.0, z)
    out = {}
    for tie in (0, 1):
        out[tie] = oracle.fit_evaluate_paths(plan, 110.0, 100.0, training, evaluation, degree=0,
                                             tie_rule=tie, tie_window=0.0)
        assert out[tie]["fits"][0]["kind"] == "regression"
        # every path is in the money at the exercise node and its continuation is the mean of
        # identical responses, i.e. its own intrinsic value
        assert out[tie]["near_tie_training_decisions"] == n
        assert out[tie]["near_tie_evaluation_decisions"] == n
        assert out[tie]["price"] == 10.0
    assert out[0]["exercise_counts"] == [0, n]   # strictly greater: nobody exercises at a tie
    assert out[1]["exercise_counts"] == [n, 0]   # greater or equal: everybody does


def _recurrence(model: dict, x: np.ndarray) -> np.ndarray:
    """Continuation of a fitted rule: Stieltjes basis by its three-term recurrence."""
    z = (x - model["mean"]) / model["scale"] if model["degree"] > 0 else np.zeros_like(x)
    p_prev = np.zeros_like(x)
    p = np.ones_like(x)
    total = model["coef"][0] * p
    for k in range(model["degree"]):
        p_next = (z - model["alpha"][k]) * p - (model["beta"][k] * p_prev if k > 0 else 0.0)
        p_prev, p = p, p_next
        total = total + model["coef"][k + 1] * p
    return total


def _near(intrinsic, continuation, window):
    return np.abs(intrinsic - continuation) <= window * np.abs(continuation)


@pytest.mark.parametrize("window", [2.0 ** -6, 2.0 ** -10])
def test_near_tie_counts_equal_an_independent_recount(oracle, window):
    times = [0.0, 0.25, 0.5, 0.75, 1.0]
    r, q = 0.03, 0.06
    plan = oracle.Plan(times=times, log_discounts=[-r * t for t in times],
                       cum_yields=[q * t for t in times], volatility=0.3,
                       exercise_nodes=[1, 2, 3], expiry_node=4)
    rng = np.random.default_rng(7)
    n, strike, spot = 4000, 95.0, 100.0
    training = oracle.simulate_normals(plan, spot, rng.standard_normal((4, n)))
    evaluation = oracle.simulate_normals(plan, spot, rng.standard_normal((4, n)))
    res = oracle.fit_evaluate_paths(plan, spot, strike, training, evaluation, tie_window=window)
    fits = res["fits"]
    # training: every regression rule is applied once to the in-the-money paths of its row
    for row, f in enumerate(fits):
        s = training[row]
        itm = s > strike
        expected = 0
        if f["kind"] == "regression":
            expected = int(np.count_nonzero(_near(s[itm] - strike, _recurrence(f["model"], s[itm] / strike), window)))
        assert f["near_ties"] == expected, (row, f["near_ties"], expected)
    assert res["near_tie_training_decisions"] == sum(f["near_ties"] for f in fits)
    # evaluation: rows in order, active in-the-money paths, stop where intrinsic > continuation
    active = np.ones(n, dtype=bool)
    ties = 0
    counts = []
    for row, f in enumerate(fits):
        s = evaluation[row]
        cand = active & (s > strike)
        stopped = 0
        if f["kind"] == "regression" and cand.any():
            cont = _recurrence(f["model"], s[cand] / strike)
            ties += int(np.count_nonzero(_near(s[cand] - strike, cont, window)))
            ex = (s[cand] - strike) > cont
            idx = np.flatnonzero(cand)[ex]
            active[idx] = False
            stopped = idx.size
        counts.append(stopped)
    assert counts == res["exercise_counts"][:-1]
    assert res["near_tie_evaluation_decisions"] == ties
    assert ties > 0 and res["near_tie_training_decisions"] > 0


def test_the_window_changes_only_the_counts(oracle):
    times = [0.0, 0.5, 1.0]
    plan = oracle.Plan(times=times, log_discounts=[-0.02 * t for t in times],
                       cum_yields=[0.05 * t for t in times], volatility=0.25,
                       exercise_nodes=[1], expiry_node=2)
    rng = np.random.default_rng(3)
    training = oracle.simulate_normals(plan, 100.0, rng.standard_normal((2, 2000)))
    evaluation = oracle.simulate_normals(plan, 100.0, rng.standard_normal((2, 2000)))
    narrow = oracle.fit_evaluate_paths(plan, 100.0, 90.0, training, evaluation, tie_window=0.0)
    wide = oracle.fit_evaluate_paths(plan, 100.0, 90.0, training, evaluation, tie_window=0.5)
    assert narrow["near_tie_training_decisions"] == narrow["near_tie_evaluation_decisions"] == 0
    assert wide["near_tie_training_decisions"] > 0 and wide["near_tie_evaluation_decisions"] > 0
    for key in ("price", "sd", "se", "exercise_counts"):
        assert narrow[key] == wide[key], key
