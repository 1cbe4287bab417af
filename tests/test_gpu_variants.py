"""Reference variants on the GPU against the oracle: the cash-dividend-date exercise mesh and its
rejections, the survivor cash policy, the GREATER_OR_EQUAL tie rule, the AT_OR_BEFORE maturity
boundary with a terminal dividend, and the reference's bitwise mesh invariants run through the HIP
path (crates/finoptions-lsm/src/batch/pricing.rs:751-838,1338-1390;
tests/test_batch_lsm.py:488-740; crates/finoptions-core/src/dividend.rs:59-81;
crates/finoptions-lsm/src/policy.rs:21-26).

Bars: FP64 prices, standard errors and deviations within 1e-10 relative of the oracle, every
count column equal; the fused method within 0.01 SE; HIP-vs-HIP invariants bitwise.
"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import rel_err

pytestmark = pytest.mark.gpu

TOL = 1e-10
COUNTS = ("origin_decisions_fitted", "origin_exercised", "regression_nodes",
          "no_training_support_nodes", "insufficient_exercise_support_nodes",
          "minimum_in_the_money_training_paths", "early_exercise_paths", "expiry_paths")
FLOATS = ("prices", "sample_standard_deviations", "standard_errors",
          "maximum_regression_condition_numbers")


def _spec(**kw):
    spec = dict(spots=(100.0,), strikes=(100.0,), maturities=(1.0,), rates=(0.03,), vols=(0.25,),
                yields=None, cash=None, policy=0, boundary=0, mesh=("uniform", 8, 0), paths=8192,
                seeds=(7, 11), domains=(1401, 91401), tie=0, support=0, offsets=None)
    spec.update(kw)
    return spec


def _hip(fh, s, executor, *, row_batch=True, greeks=()):
    n = len(s["spots"])
    offsets = s["offsets"] if s["offsets"] is not None else np.arange(n + 1, dtype=np.uint32)
    if s["cash"] is not None:
        co, ct, ca = s["cash"]
        dividends = fh.CashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=np.array(co, dtype=np.int64),
                                             times=np.array(ct, dtype=np.float64),
                                             amounts=np.array(ca, dtype=np.float64)),
            policy=fh.DividendPolicy.LIQUIDATOR if s["policy"] == 0 else fh.DividendPolicy.SURVIVOR,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE if s["boundary"] == 0
            else fh.MaturityBoundary.AT_OR_BEFORE)
    else:
        dividends = fh.ContinuousDividendYieldBatch(
            annualized_decimals=np.array(s["yields"] or [0.0] * n, dtype=np.float64))
    kind = s["mesh"][0]
    origin = fh.OriginExercise.EXCLUDED if s["mesh"][-1] == 0 else fh.OriginExercise.INCLUDED
    if kind == "uniform":
        mesh = fh.UniformExerciseMesh(base_intervals=s["mesh"][1], origin_exercise=origin)
    elif kind == "cash":
        mesh = fh.CashDividendExerciseMesh()
    else:
        eo, et = s["mesh"][1], s["mesh"][2]
        mesh = fh.ExplicitExerciseMesh(
            schedules=fh.RaggedExerciseTimes(offsets=np.array(eo, dtype=np.int64),
                                             times=np.array(et, dtype=np.float64)),
            origin_exercise=origin)
    method = fh.Lsm(training_paths=s["paths"], evaluation_paths=s["paths"],
                    training_stream=fh.RandomStream(seed=s["seeds"][0], domain=s["domains"][0]),
                    evaluation_stream=fh.RandomStream(seed=s["seeds"][1], domain=s["domains"][1]),
                    insufficient_exercise_support=(
                        fh.InsufficientExerciseSupportPolicy.FAIL if s["support"] == 0
                        else fh.InsufficientExerciseSupportPolicy.CONTINUE),
                    tie_rule=fh.ExerciseTieRule.STRICTLY_GREATER if s["tie"] == 0
                    else fh.ExerciseTieRule.GREATER_OR_EQUAL)
    f64 = lambda v: np.array(v, dtype=np.float64)
    if row_batch and s["offsets"] is None:
        return fh.price_lsm_call_batch(initial_spots=f64(s["spots"]), strikes=f64(s["strikes"]),
                                       maturities=f64(s["maturities"]), rates=f64(s["rates"]),
                                       volatilities=f64(s["vols"]), dividends=dividends, mesh=mesh,
                                       executor=executor, method=method, greeks=greeks)
    groups = fh.LsmStrikeGroups(group_offsets=np.asarray(offsets, dtype=np.uint32),
                                initial_spots=f64(s["spots"]), strikes=f64(s["strikes"]),
                                maturities=f64(s["maturities"]), rates=f64(s["rates"]),
                                volatilities=f64(s["vols"]), dividends=dividends)
    return fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


def _oracle(oracle, s):
    n = len(s["spots"])
    offsets = s["offsets"] if s["offsets"] is not None else np.arange(n + 1, dtype=np.uint32)
    cols = dict(group_offsets=np.asarray(offsets, dtype=np.uint32), initial_spots=np.array(s["spots"]),
                strikes=np.array(s["strikes"]), maturities=np.array(s["maturities"]),
                rates=np.array(s["rates"]), volatilities=np.array(s["vols"]))
    if s["cash"] is not None:
        co, ct, ca = s["cash"]
        cols.update(cash_offsets=np.array(co, dtype=np.int64), cash_times=np.array(ct),
                    cash_amounts=np.array(ca), cash_policy=s["policy"])
    else:
        cols["dividend_yields"] = np.array(s["yields"] or [0.0] * n)
    kind = s["mesh"][0]
    if kind == "uniform":
        mesh = dict(mesh_kind=0, base_intervals=s["mesh"][1], origin_exercise=s["mesh"][-1])
    elif kind == "cash":
        mesh = dict(mesh_kind=1)
    else:
        mesh = dict(mesh_kind=2, exercise_offsets=np.array(s["mesh"][1], dtype=np.int64),
                    exercise_times=np.array(s["mesh"][2]), origin_exercise=s["mesh"][-1])
    req = oracle.GroupedRequest(training_paths=s["paths"], evaluation_paths=s["paths"],
                                training_seed=s["seeds"][0], evaluation_seed=s["seeds"][1],
                                training_domain_start=s["domains"][0],
                                evaluation_domain_start=s["domains"][1], tie_rule=s["tie"],
                                insufficient_support=s["support"], **cols, **mesh)
    return oracle.price_grouped(req), req


def _assert_matches_oracle(res, ref, label):
    for name in FLOATS:
        err = rel_err(getattr(res, name), ref[name])
        assert err < TOL, (label, name, err)
    for name in COUNTS:
        assert np.array_equal(getattr(res, name), ref[name]), (label, name)
    print(f"{label}: max price rel err {rel_err(res.prices, ref['prices']):.2e}")


def _assert_bitwise(a, b, label):
    for name in FLOATS + COUNTS:
        x, y = np.asarray(getattr(a, name)), np.asarray(getattr(b, name))
        if x.dtype == np.float64:
            x, y = x.view(np.uint64), y.view(np.uint64)
        assert np.array_equal(x, y), (label, name)


# ------------------------------------------------------------------ cash-dividend-date mesh

def test_cash_dividend_date_mesh_matches_the_oracle_and_the_uniform_mesh(hip_f64, hip_fused, oracle):
    import finoptions_hip as fh

    # tests/test_batch_lsm.py:488-520 geometry, then a two-dividend, two-group request
    one = _spec(strikes=(95.0,), cash=([0, 1], [0.5], [2.0]), mesh=("cash", 0),
                domains=(913, 90913))
    special = _hip(fh, one, hip_f64)
    uniform = _hip(fh, dict(one, mesh=("uniform", 2, 0)), hip_f64)
    _assert_bitwise(special, uniform, "cash-date == uniform(2)")
    assert special.regression_nodes[0] == 1
    ref, _ = _oracle(oracle, one)
    _assert_matches_oracle(special, ref, "cash-date mesh, one dividend")
    two = _spec(spots=(100.0, 80.0), strikes=(90.0, 85.0, 100.0), offsets=[0, 2, 3],
                maturities=(1.0, 1.5), rates=(0.03, 0.0), vols=(0.3, 0.2),
                cash=([0, 2, 3], [0.3, 0.6, 0.9], [3.0, 2.0, 4.0]), mesh=("cash", 0), paths=16384)
    ref, _ = _oracle(oracle, two)
    for hip, label in ((hip_f64, "F64"), (hip_fused, "fused")):
        res = _hip(fh, two, hip, row_batch=False)
        if label == "F64":
            _assert_matches_oracle(res, ref, "cash-date mesh, two groups")
        else:
            bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
            assert bias.max() < 0.01, bias


def test_cash_dividend_date_mesh_rejects_every_out_of_domain_request(hip_f64):
    """tests/test_batch_lsm.py:690-740 with the reference's messages (batch/error.rs:81-100)."""
    import finoptions_hip as fh

    base = _spec(mesh=("cash", 0), domains=(1301, 91301))
    with pytest.raises(fh.InputError, match="requires cash-dividend inputs"):
        _hip(fh, dict(base, yields=[0.0]), hip_f64)
    with pytest.raises(fh.InputError, match=r"requires a nonnegative.*rate.*row 0; observed -0\.001"):
        _hip(fh, dict(base, rates=(-0.001,), cash=([0, 1], [0.5], [2.0])), hip_f64)
    with pytest.raises(fh.InputError,
                       match=r"does not admit terminal cash.*row 0, event 0; event and maturity are 1\.0"):
        _hip(fh, dict(base, cash=([0, 1], [1.0], [2.0]), boundary=1), hip_f64)
    # the explicit mesh rejects an opportunity at maturity (tests/test_batch_lsm.py:587-624)
    with pytest.raises(fh.InputError, match=r"must be strictly before maturity.*1\.0"):
        _hip(fh, _spec(mesh=("explicit", [0, 1], [1.0], 0), paths=512), hip_f64)


# ------------------------------------------------------------------------ survivor policy

def test_survivor_cash_policy_matches_the_oracle(hip_f64, hip_fused, oracle):
    """dividend.rs:70-76: a dividend larger than the spot leaves the survivor's spot unchanged and
    the liquidator's at zero; both policies against the oracle on the same request."""
    import finoptions_hip as fh

    spec = _spec(spots=(100.0,), strikes=(20.0, 40.0, 60.0, 95.0), offsets=[0, 4], vols=(0.6,),
                 cash=([0, 2], [0.3, 0.6], [60.0, 10.0]), mesh=("uniform", 10, 0), paths=16384,
                 support=1)
    prices = {}
    for policy in (1, 0):
        s = dict(spec, policy=policy)
        ref, _ = _oracle(oracle, s)
        res = _hip(fh, s, hip_f64, row_batch=False)
        _assert_matches_oracle(res, ref, f"cash policy {policy}")
        fused = _hip(fh, s, hip_fused, row_batch=False)
        bias = np.abs(fused.prices - ref["prices"]) / ref["standard_errors"]
        assert bias.max() < 0.01, bias
        prices[policy] = res.prices
    assert np.all(prices[1] > prices[0])   # surviving spots keep value the liquidator loses


def test_survivor_greeks_match_the_oracle(hip_f64, oracle):
    """Survivor tangents pass through a dividend whether or not it exceeds the spot."""
    import finoptions_hip as fh

    s = _spec(spots=(100.0,), strikes=(30.0, 60.0, 100.0), offsets=[0, 3], vols=(0.6,),
              cash=([0, 1], [0.5], [70.0]), policy=1, mesh=("uniform", 6, 0), paths=16384,
              support=1)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO)
    res = _hip(fh, s, hip_f64, row_batch=False, greeks=greeks)
    _, req = _oracle(oracle, s)
    ref = oracle.grouped_pathwise_greeks(req, 0)
    for g in greeks:
        err = rel_err(res.greeks[g], ref[g.value])
        print(f"survivor {g.value}: rel err {err:.2e}")
        assert err < 1e-9


# ----------------------------------------------------------------------------- tie rule

def test_greater_or_equal_tie_rule_matches_the_oracle(hip_f64, hip_fused, oracle):
    """policy.rs:21-26 GreaterOrEqual on the device and on the oracle, and its exact-tie counts."""
    import finoptions_hip as fh

    spec = _spec(spots=(100.0,), strikes=tuple(np.linspace(85.0, 115.0, 12)), offsets=[0, 12],
                 yields=[0.02], mesh=("uniform", 20, 0), paths=32768, tie=1, support=1)
    ref, _ = _oracle(oracle, spec)
    res = _hip(fh, spec, hip_f64, row_batch=False)
    _assert_matches_oracle(res, ref, "GREATER_OR_EQUAL F64")
    assert np.array_equal(res.near_tie_training_decisions, ref["near_tie_training_decisions"])
    assert np.array_equal(res.near_tie_evaluation_decisions, ref["near_tie_evaluation_decisions"])
    fused = _hip(fh, spec, hip_fused, row_batch=False)
    bias = np.abs(fused.prices - ref["prices"]) / ref["standard_errors"]
    assert bias.max() < 0.01
    strict = _hip(fh, dict(spec, tie=0), hip_f64, row_batch=False)
    ties = int(ref["near_tie_training_decisions"].sum() + ref["near_tie_evaluation_decisions"].sum())
    if ties == 0:   # without exact ties the two rules decide identically
        _assert_bitwise(strict, res, "tie rules without ties")


def test_tie_rules_split_exact_ties_at_the_origin(hip_f64, oracle):
    """An origin exercise opportunity whose continuation estimate equals the intrinsic value
    exactly: a dividend that liquidates every path before the only later exercise date makes the
    continuation 0, and a strike equal to the spot makes the intrinsic 0. STRICTLY_GREATER holds,
    GREATER_OR_EQUAL exercises at the origin (policy.rs:21-26, 267-281)."""
    import finoptions_hip as fh

    spec = _spec(spots=(50.0,), strikes=(50.0,), cash=([0, 1], [0.25], [1000.0]),
                 mesh=("explicit", [0, 1], [0.5], 1), paths=4096, support=1)
    for tie in (0, 1):
        s = dict(spec, tie=tie)
        res = _hip(fh, s, hip_f64)
        ref, _ = _oracle(oracle, s)
        _assert_matches_oracle(res, ref, f"origin tie, rule {tie}")
        assert res.origin_decisions_fitted[0]
        assert bool(res.origin_exercised[0]) == (tie == 1)
        assert res.prices[0] == 0.0
        assert res.early_exercise_paths[0] == (4096 if tie == 1 else 0)


# -------------------------------------------------------------- AT_OR_BEFORE terminal cash

def test_at_or_before_terminal_cash_explicit_equals_uniform(hip_f64, hip_fused, oracle):
    """pricing.rs:1338-1390 on the device: cash 4.0 at maturity, uniform(4) and explicit
    [0.25, 0.5, 0.75] are bitwise equal, with three regression nodes, and equal the oracle."""
    import finoptions_hip as fh

    common = dict(cash=([0, 1], [1.0], [4.0]), boundary=1, paths=2048, domains=(1401, 91401))
    uniform_s = _spec(mesh=("uniform", 4, 0), **common)
    exact_s = _spec(mesh=("explicit", [0, 3], [0.25, 0.5, 0.75], 0), **common)
    for hip in (hip_f64, hip_fused):
        uniform = _hip(fh, uniform_s, hip)
        exact = _hip(fh, exact_s, hip)
        _assert_bitwise(exact, uniform, "explicit == uniform, terminal cash")
        assert exact.regression_nodes[0] == 3
    ref, _ = _oracle(oracle, exact_s)
    _assert_matches_oracle(_hip(fh, exact_s, hip_f64), ref, "AT_OR_BEFORE explicit")
    # STRICTLY_BEFORE rejects the same terminal dividend
    with pytest.raises(fh.InputError):
        _hip(fh, dict(uniform_s, boundary=0), hip_f64)


def test_at_or_before_sparse_terminal_predecessor(hip_f64, oracle):
    """tests/test_batch_lsm.py:565-584: exercise at 255/256 only, terminal dividend at 1.0."""
    import finoptions_hip as fh

    s = _spec(mesh=("explicit", [0, 1], [255.0 / 256.0], 0), cash=([0, 1], [1.0], [4.0]),
              boundary=1, paths=2048)
    res = _hip(fh, s, hip_f64)
    assert res.regression_nodes[0] == 1
    ref, _ = _oracle(oracle, s)
    _assert_matches_oracle(res, ref, "sparse terminal predecessor")


def test_at_or_before_terminal_cash_grouped_strikes_match_the_oracle(hip_f64, hip_fused, oracle):
    """Grouped strikes with quarterly dividends, the last one at maturity, on a uniform mesh, both
    policies; the survivor keeps spots the terminal dividend exceeds."""
    import finoptions_hip as fh

    for policy in (0, 1):
        s = _spec(spots=(100.0,), strikes=tuple(np.linspace(80.0, 120.0, 10)), offsets=[0, 10],
                  cash=([0, 4], [0.25, 0.5, 0.75, 1.0], [1.0, 1.5, 2.0, 30.0]), boundary=1,
                  policy=policy, mesh=("uniform", 12, 0), paths=16384, support=1)
        ref, _ = _oracle(oracle, s)
        _assert_matches_oracle(_hip(fh, s, hip_f64, row_batch=False), ref,
                               f"AT_OR_BEFORE grouped, policy {policy}")
        fused = _hip(fh, s, hip_fused, row_batch=False)
        assert (np.abs(fused.prices - ref["prices"]) / ref["standard_errors"]).max() < 0.01


# ------------------------------------------------------------- explicit mesh: rows, groups

def test_explicit_mesh_is_identical_for_row_and_grouped_batches(hip_f64, hip_fused, oracle):
    """tests/test_batch_lsm.py:627-688: explicit [0.5], [0.375] equal uniform(2) for maturities
    1.0 and 0.75, through the row batch and the grouped batch."""
    import finoptions_hip as fh

    common = dict(spots=(100.0, 120.0), strikes=(100.0, 110.0), maturities=(1.0, 0.75),
                  rates=(0.05, 0.03), vols=(0.25, 0.30), yields=[0.02, 0.01], paths=2048,
                  domains=(1601, 91601))
    explicit = _spec(mesh=("explicit", [0, 1, 2], [0.5, 0.375], 0), **common)
    uniform = _spec(mesh=("uniform", 2, 0), **common)
    for hip in (hip_f64, hip_fused):
        rows_e = _hip(fh, explicit, hip)
        _assert_bitwise(rows_e, _hip(fh, uniform, hip), "row explicit == uniform")
        grouped_e = _hip(fh, dict(explicit, offsets=[0, 1, 2]), hip, row_batch=False)
        _assert_bitwise(grouped_e, rows_e, "grouped == row")
    ref, _ = _oracle(oracle, explicit)
    _assert_matches_oracle(_hip(fh, explicit, hip_f64), ref, "explicit rows")
