"""GPU parity of the HIP engine against the CPU oracle on identical seeded inputs.

Bars: Philox words bit-exact; ZIGNOR normals bit-exact wherever the accept decision needs no
transcendental (every normal drawn with one u64); FP64 paths within a few ulps; plain FP64 prices,
standard errors, regression coefficients within 1e-10 relative; stopping counts equal. Flips of
exercise decisions at ties are counted and reported.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import (
    c1_inputs,
    c1_plan,
    c2_columns,
    hip_groups,
    hip_method,
    oracle_grouped,
    rel_err,
)

pytestmark = pytest.mark.gpu

PRICE_TOL = 1e-10


def test_philox_words_are_bit_exact(hip_f64, oracle, golden):
    import finoptions_hip.simulation as sim

    for kat in golden["philox_kats"]:
        if kat["counter"][1] != 0:
            continue
        seed = kat["key"][0] | (kat["key"][1] << 32)
        words = sim.philox_blocks(hip_f64, seed=seed, domain=kat["counter"][3],
                                  path0=kat["counter"][2], paths=1, blocks=1)
        assert words[0, 0].tolist() == kat["expected"]
    for seed, domain in ((7, 1), (0x0123456789ABCDEF, 0xFFFFFFFF), (2**64 - 1, 0)):
        words = sim.philox_blocks(hip_f64, seed=seed, domain=domain, path0=1000, paths=64,
                                  blocks=16)
        for p in (0, 1, 31, 63):
            for b in (0, 1, 15):
                ref = oracle.philox([b, 0, 1000 + p, domain], [seed & 0xFFFFFFFF, seed >> 32])
                assert words[p, b].tolist() == ref.tolist()


def test_normals_match_the_oracle_stream(hip_f64, oracle, golden):
    import finoptions_hip.simulation as sim

    z = golden["zignor_sequence"]
    values, draws = sim.stream_normals(hip_f64, seed=z["seed"], domain=z["domain"],
                                       path0=z["path"], paths=1, count=len(z["bits"]))
    assert [f"0x{v:016x}" for v in values[0].view(np.uint64)] == z["bits"]
    paths, count = 512, 96
    dev, dev_draws = sim.stream_normals(hip_f64, seed=11, domain=1, path0=0, paths=paths,
                                        count=count)
    flips = 0
    ulp_diff = 0
    for p in range(paths):
        ref, ref_draws = oracle.normals(11, 1, p, count)
        same_draws = np.array_equal(ref_draws, dev_draws[p])
        if not same_draws:
            flips += 1       # a wedge/tail accept decision went the other way (exp/log ulp)
            continue
        one = ref_draws == 1
        assert np.array_equal(ref[one].view(np.uint64), dev[p][one].view(np.uint64))
        rest = ~one
        ulp_diff += int(np.sum(ref[rest].view(np.uint64) != dev[p][rest].view(np.uint64)))
        np.testing.assert_allclose(dev[p][rest], ref[rest], rtol=1e-14)
    print(f"normal stream: {paths * count} normals, decision flips on {flips} paths, "
          f"{ulp_diff} multi-draw normals differing in the last bits")
    assert flips <= 1


def test_squeezed_wedge_tests_match_the_oracle_stream(hip_f64, oracle):
    """2^21 normals (about 31,000 wedge tests, most decided by the layer squeeze, a few dozen by
    the exponential inside its band): every draw count equal to the oracle's exact stream,
    single-draw normals bit for bit, wedge-accepted normals too (they are value = unit * X)."""
    import finoptions_hip.simulation as sim

    paths, count = 16384, 128
    dev, dev_draws = sim.stream_normals(hip_f64, seed=2024, domain=7, path0=0, paths=paths,
                                        count=count)
    mismatched = 0
    wedge = 0
    for p in range(paths):
        ref, ref_draws = oracle.normals(2024, 7, p, count)
        if not np.array_equal(ref_draws, dev_draws[p]):
            mismatched += 1
            continue
        wedge += int(np.sum(ref_draws > 1))
        # every normal but a tail one is unit * X[index], the same product on both sides
        same = np.abs(ref) < 3.6541528853610088
        assert np.array_equal(ref[same].view(np.uint64), dev[p][same].view(np.uint64))
        np.testing.assert_allclose(dev[p][~same], ref[~same], rtol=1e-14)
    print(f"squeezed stream: {paths * count} normals, {wedge} multi-draw normals, "
          f"{mismatched} paths with a different draw count")
    assert wedge > 20_000
    assert mismatched <= 1   # a tail decision may flip on the device libm's log (last bit)


def test_fp64_paths_match_the_oracle(hip_f64, oracle):
    import finoptions_hip as fh

    inp = c1_inputs()
    paths = fh.simulate_gbm_seeded(initial_spot=100.0, schedule=inp["schedule"], curve=inp["curve"],
                                   dividend_yield_curve=inp["dividend_yield_curve"],
                                   volatility=inp["volatility"],
                                   dividend_policy=inp["dividend_policy"], paths=65_536,
                                   random_stream=fh.RandomStream(seed=7, domain=1),
                                   executor=hip_f64)
    ref = oracle.simulate_seeded(c1_plan(), 100.0, 65_536, 7, 1)
    assert paths.states.shape == ref.shape
    err = rel_err(paths.states, ref)
    exact = float(np.mean(paths.states == ref))
    print(f"FP64 paths: max rel err {err:.3e}, bit-identical fraction {exact:.4f}")
    assert err < 1e-13


def test_gbm_exponential_is_within_one_ulp_of_the_c_library(hip_f64):
    """The device's table-driven exp of the GBM step against the C library's (Python's math.exp):
    at most one ulp apart on random arguments of the step's range and on wide ones, equal on the
    large majority, and equal on the special values the step can produce."""
    import math

    import finoptions_hip as fh

    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-5.0, 5.0, 200_000), rng.uniform(-1e-3, 1e-3, 50_000),
                        rng.uniform(-700.0, 700.0, 50_000)])
    got = fh.simulation.gbm_exp(hip_f64, x)
    ref = np.array([math.exp(v) for v in x])
    ulps = np.abs(got.view(np.int64) - ref.view(np.int64))
    print(f"table exp: max {ulps.max()} ulp, equal fraction {np.mean(ulps == 0):.4f}")
    assert ulps.max() <= 1
    assert np.mean(ulps == 0) > 0.99
    special = np.array([0.0, -0.0, 1e-300, -1e-300, 700.0, -700.0, 709.5, -745.2, 710.0, -800.0,
                        np.inf, -np.inf])
    sg = fh.simulation.gbm_exp(hip_f64, special)
    for v, g in zip(special, sg):
        r = math.exp(v) if v < 709.79 else math.inf
        assert g == r or abs(np.float64(g).view(np.int64) - np.float64(r).view(np.int64)) <= 1, (v, g, r)
    assert np.isnan(fh.simulation.gbm_exp(hip_f64, np.array([np.nan]))[0])


def test_device_tensor_export_is_the_host_export(hip_f64):
    """The observation matrix left in HBM as a torch tensor is bitwise the host export (binary64),
    or its single rounding (binary32); torch work queued before the call is ordered first."""
    import finoptions_hip as fh
    import torch

    inp = c1_inputs()
    kw = dict(initial_spot=100.0, schedule=inp["schedule"], curve=inp["curve"],
              dividend_yield_curve=inp["dividend_yield_curve"], volatility=inp["volatility"],
              dividend_policy=inp["dividend_policy"], paths=40_000,
              random_stream=fh.RandomStream(seed=7, domain=1), executor=hip_f64)
    host = fh.simulate_gbm_seeded(**kw)
    dev = fh.simulate_gbm_seeded_tensor(**kw)
    assert dev.states.device.type == "cuda" and dev.states.dtype == torch.float64
    assert dev.observation_nodes == host.observation_nodes
    assert np.array_equal(dev.states.cpu().numpy(), host.states)
    single = fh.simulate_gbm_seeded_tensor(**kw, dtype="float32")
    assert np.array_equal(single.states.cpu().numpy(), host.states.astype(np.float32))
    assert torch.equal(dev.values_at(2), dev.states[1])
    with pytest.raises(fh.InputError):
        fh.simulate_gbm_seeded_tensor(**kw, dtype="int8")


def test_c1_plain_fp64_matches_the_oracle(hip_f64, oracle):
    import finoptions_hip as fh

    method = fh.Lsm(training_paths=65_536, evaluation_paths=65_536,
                    training_stream=fh.RandomStream(seed=7, domain=1),
                    evaluation_stream=fh.RandomStream(seed=7, domain=2),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL)
    res = fh.price_lsm_call(**c1_inputs(), executor=hip_f64, method=method)
    ref = oracle.price_lsm_call(c1_plan(), 100.0, 100.0, 65_536, 65_536, 7, 1, 7, 2)
    print(f"C1 FP64: price {res.price!r} oracle {ref['price']!r} "
          f"rel {abs(res.price - ref['price']) / ref['price']:.2e}")
    assert rel_err(res.price, ref["price"]) < PRICE_TOL
    assert rel_err(res.sample_standard_deviation, ref["sd"]) < PRICE_TOL
    assert rel_err(res.standard_error, ref["se"]) < PRICE_TOL
    assert res.observation_nodes == tuple(ref["observation_nodes"])
    assert sum(res.exercise_counts) == 65_536
    flips = sum(abs(a - b) for a, b in zip(res.exercise_counts, ref["exercise_counts"])) // 2
    assert flips == 0
    d = res.fit_diagnostics[0]
    model = ref["fits"][0]["model"]
    assert isinstance(d, fh.RegressionFitDiagnostics)
    assert d.in_the_money_paths == ref["fits"][0]["itm_paths"]
    assert rel_err(d.predictor_mean, model["mean"]) < PRICE_TOL
    assert rel_err(d.predictor_scale, model["scale"]) < PRICE_TOL
    assert rel_err(d.condition_number, model["condition"]) < PRICE_TOL
    assert rel_err(res.regression_coefficients[0], model["coef"]) < PRICE_TOL


def _grouped_pair(hip, oracle, cols, *, base_intervals, paths, seed, continue_support=True,
                  eval_domain=2, **kw):
    """Price one request on the GPU and on the oracle. Deep out-of-the-money strikes at the first
    nodes have fewer in-the-money training paths than basis columns, so these requests declare
    the CONTINUE support policy (the reference requires the policy to be explicit)."""
    import finoptions_hip as fh

    groups = hip_groups(cols)
    mesh = fh.UniformExerciseMesh(base_intervals=base_intervals,
                                  origin_exercise=fh.OriginExercise.EXCLUDED)
    support = (fh.InsufficientExerciseSupportPolicy.CONTINUE if continue_support
               else fh.InsufficientExerciseSupportPolicy.FAIL)
    method = hip_method(paths, paths, seed, support=support, eval_domain=eval_domain)
    res = fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=hip, method=method,
                                          **kw)
    ref = oracle.price_grouped(oracle_grouped(cols, base_intervals=base_intervals,
                                              training_paths=paths, evaluation_paths=paths,
                                              seed=seed, eval_domain=eval_domain,
                                              insufficient_support=1 if continue_support else 0))
    return res, ref, groups, mesh, method


def test_grouped_strikes_plain_fp64_match_the_oracle(hip_f64, oracle):
    cols = c2_columns(strikes=64)
    res, ref, *_ = _grouped_pair(hip_f64, oracle, cols, base_intervals=50, paths=32_768, seed=11)
    err = rel_err(res.prices, ref["prices"])
    print(f"grouped 64 strikes x 50 intervals x 32768: max price rel err {err:.2e}")
    assert err < PRICE_TOL
    assert rel_err(res.standard_errors, ref["standard_errors"]) < PRICE_TOL
    assert rel_err(res.sample_standard_deviations, ref["sample_standard_deviations"]) < PRICE_TOL
    for name in ("regression_nodes", "no_training_support_nodes",
                 "insufficient_exercise_support_nodes", "minimum_in_the_money_training_paths",
                 "early_exercise_paths", "expiry_paths", "origin_decisions_fitted",
                 "origin_exercised"):
        assert np.array_equal(getattr(res, name), ref[name]), name
    assert rel_err(res.maximum_regression_condition_numbers,
                   ref["maximum_regression_condition_numbers"]) < PRICE_TOL


def _continuation(model_vals, x):
    """A rule's continuation at moneyness x (mean, scale, alpha[4], beta[4], coef[5]): the
    three-term recurrence of the orthogonal basis, evaluated in numpy for both rules alike."""
    mean, scale = model_vals[0], model_vals[1]
    alpha, beta, coef = model_vals[2:6], model_vals[6:10], model_vals[10:15]
    z = (x - mean) / scale
    pred, prev, cur = np.full_like(x, coef[0]), np.zeros_like(x), np.ones_like(x)
    for k in range(4):
        nxt = (z - alpha[k]) * cur - beta[k] * prev
        pred = pred + coef[k + 1] * nxt
        prev, cur = cur, nxt
    return pred


COEF_CASES = {
    "8 strikes x 12 rows, yield, 2^14 paths": dict(strikes=8, intervals=12, paths=16_384, cash=False),
    "16 strikes x 12 rows, cash dividends, 2^14 paths": dict(strikes=16, intervals=12, paths=16_384, cash=True),
    "64 strikes x 50 rows, yield, 2^13 paths": dict(strikes=64, intervals=50, paths=8_192, cash=False),
}


@pytest.mark.parametrize("case", list(COEF_CASES))
def test_grouped_coefficients_match_the_oracle(hip_f64, oracle, case):
    """Every fitted rule of the FP64 method against the oracle's: each coefficient field within
    1e-10 of its natural scale (the largest value the field takes over the group's rules, at least
    1 for the recurrence coefficients of the standardized moneyness), and the rule's continuation
    on a grid of standardized moneyness z in [-3, 3] within 1e-10 of the rule's largest
    continuation there (the quantity exercise decisions compare)."""
    import finoptions_hip as fh

    c = COEF_CASES[case]
    cols = c2_columns(strikes=c["strikes"])
    if c["cash"]:
        del cols["yields"]
        cols.update(cash_offsets=np.array([0, 3], dtype=np.int64),
                    cash_times=np.array([0.25, 0.5, 0.75]), cash_amounts=np.array([0.6, 0.8, 1.0]))
    groups = hip_groups(cols)
    mesh = fh.UniformExerciseMesh(base_intervals=c["intervals"], origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(c["paths"], c["paths"], 5, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    coef = fh.grouped_fit_coefficients(groups=groups, mesh=mesh, executor=hip_f64, method=method,
                                       group=0)
    req = oracle_grouped(cols, base_intervals=c["intervals"], training_paths=c["paths"],
                         evaluation_paths=c["paths"], seed=5, insufficient_support=1)
    models = oracle.grouped_models(req, 0, coef.shape[1])
    got, ref = [], []
    for s in range(coef.shape[0]):
        for r in range(coef.shape[1]):
            if models[s][r]["kind"] != "regression":
                continue
            m = models[s][r]["model"]
            ref.append([m["mean"], m["scale"], *m["alpha"], *m["beta"], *m["coef"], m["condition"]])
            got.append(coef[s, r, :16])
    got, ref = np.array(got), np.array(ref)
    assert len(ref) > 0
    names = ["mean", "scale"] + [f"alpha{k}" for k in range(4)] + [f"beta{k}" for k in range(4)] + \
        [f"coef{k}" for k in range(5)] + ["condition"]
    # natural scale of each field: its largest value over the rules; the recurrence coefficients
    # alpha and beta of the standardized moneyness are O(1) numbers (alpha0 is exactly 0 in exact
    # arithmetic, so its values are rounding noise), so theirs is at least 1
    field_scale = np.max(np.abs(ref), axis=0)
    field_scale[2:10] = np.maximum(field_scale[2:10], 1.0)
    field_err = np.max(np.abs(got - ref), axis=0) / np.maximum(field_scale, 1e-300)
    elementwise = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3)
    z = np.linspace(-3.0, 3.0, 61)
    pred_err = 0.0
    for g, f in zip(got, ref):
        x = f[0] + z * f[1]
        pg, pf = _continuation(g, x), _continuation(f, x)
        pred_err = max(pred_err, float(np.max(np.abs(pg - pf)) / max(np.max(np.abs(pf)), 1e-12)))
    print(f"{case}: {len(ref)} rules; continuation {pred_err:.2e}; fields: " +
          ", ".join(f"{n} {e:.1e}" for n, e in zip(names, field_err)) +
          f"; elementwise with a 1e-3 floor {elementwise.max():.2e}")
    assert field_err.max() < 1e-10
    assert pred_err < 1e-10


def test_grouped_cash_dividends_match_the_oracle(hip_f64, oracle):
    cols = c2_columns(strikes=16)
    del cols["yields"]
    cols.update(cash_offsets=np.array([0, 3], dtype=np.int64),
                cash_times=np.array([0.25, 0.5, 0.75]), cash_amounts=np.array([0.6, 0.8, 1.0]))
    res, ref, *_ = _grouped_pair(hip_f64, oracle, cols, base_intervals=12, paths=16_384, seed=3)
    assert rel_err(res.prices, ref["prices"]) < PRICE_TOL
    assert np.array_equal(res.expiry_paths, ref["expiry_paths"])


def test_results_are_deterministic_and_tile_invariant(hip_f64):
    import finoptions_hip as fh

    cols = c2_columns(strikes=16)
    groups = hip_groups(cols)
    mesh = fh.UniformExerciseMesh(base_intervals=10, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(8192, 8192, 17)
    outs = []
    for tile in (0, 1, 4, 16, 0):
        hip_f64.tune(eval_strike_tile=tile)
        outs.append(fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=hip_f64,
                                                    method=method))
    hip_f64.tune(eval_strike_tile=0)
    for o in outs[1:]:
        assert np.array_equal(o.prices.view(np.uint64), outs[0].prices.view(np.uint64))
        assert np.array_equal(o.standard_errors, outs[0].standard_errors)
        assert np.array_equal(o.expiry_paths, outs[0].expiry_paths)


@pytest.mark.parametrize("intervals", [10, 120])
def test_fused_results_are_tile_invariant(hip_fused, intervals):
    """The fused evaluation has three forms: single-strike tiles regenerate the normals from the
    streams; wider tiles read normals drawn once per pricing and stage the tile's decision data in
    LDS when it fits (8 strikes x 121 rows does not, 4 x 121 does). Every form gives the same
    bits, with and without Greeks."""
    import finoptions_hip as fh

    cols = c2_columns(strikes=16)
    groups = hip_groups(cols)
    mesh = fh.UniformExerciseMesh(base_intervals=intervals, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(8192, 8192, 23, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    for greeks, tiles in (((fh.Greek.DELTA, fh.Greek.VEGA), (0, 1, 4, 8)), ((), (0, 1, 4, 16))):
        outs = []
        for tile in tiles:
            hip_fused.tune(eval_strike_tile=tile)
            outs.append(fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=hip_fused,
                                                        method=method, greeks=greeks))
        hip_fused.tune(eval_strike_tile=0)
        for tile, o in zip(tiles[1:], outs[1:]):
            assert np.array_equal(o.prices.view(np.uint64), outs[0].prices.view(np.uint64)), tile
            assert np.array_equal(o.standard_errors, outs[0].standard_errors), tile
            assert np.array_equal(o.early_exercise_paths, outs[0].early_exercise_paths), tile
            assert np.array_equal(o.expiry_paths, outs[0].expiry_paths), tile
            for g in greeks:
                assert np.array_equal(o.greeks[g], outs[0].greeks[g]), (tile, g)


def test_fused_method_matches_the_oracle(hip_fused, oracle):
    """One fused pass per node (pivoted power sums, moment-form Stieltjes solve) over binary64
    spots and cashflows: the fitted rules agree with the reference's to ~1e-13, so exercise
    decisions agree and prices match far inside the statistical error."""
    cols = c2_columns(strikes=64)
    res, ref, *_ = _grouped_pair(hip_fused, oracle, cols, base_intervals=50, paths=32_768, seed=11)
    err = rel_err(res.prices, ref["prices"])
    bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
    print(f"fused FP64 vs oracle: max price rel err {err:.2e}, max |bias|/SE {bias.max():.2e}")
    assert bias.max() < 0.01
    assert err < 1e-9
    assert np.array_equal(res.expiry_paths, ref["expiry_paths"])


def test_fp32_bulk_typical_strike_stays_within_a_hundredth_of_a_standard_error(hip_f32, oracle):
    """Binary32 spots and cashflows flip rare near-tie exercise decisions; backward induction
    amplifies a flip. The typical strike stays within 0.01 SE of the FP64 oracle."""
    cols = c2_columns(strikes=64)
    res, ref, *_ = _grouped_pair(hip_f32, oracle, cols, base_intervals=50, paths=32_768, seed=11)
    bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
    print(f"FP32 bulk vs FP64 oracle: max |bias|/SE {bias.max():.3f}, median {np.median(bias):.2e}, "
          f"strikes over 0.1 SE: {int(np.sum(bias > 0.1))}")
    assert np.median(bias) < 0.01


@pytest.mark.xfail(strict=True, reason="binary32 storage flips near-tie decisions: single strikes "
                   "of the 64-strike state move by 0.1-0.5 SE (measured 0.146 SE at 32768 paths)")
def test_fp32_bulk_meets_the_tenth_of_a_standard_error_bar(hip_f32, oracle):
    cols = c2_columns(strikes=64)
    res, ref, *_ = _grouped_pair(hip_f32, oracle, cols, base_intervals=50, paths=32_768, seed=11)
    bias = np.abs(res.prices - ref["prices"]) / ref["standard_errors"]
    assert bias.max() < 0.1


def test_pathwise_greeks_match_the_oracle(hip_f64, oracle):
    import finoptions_hip as fh

    cols = c2_columns(strikes=8)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)
    res, _ref, *_ = _grouped_pair(hip_f64, oracle, cols, base_intervals=8, paths=16_384, seed=9,
                                  greeks=greeks)
    req = oracle_grouped(cols, base_intervals=8, training_paths=16_384, evaluation_paths=16_384,
                         seed=9, insufficient_support=1)
    ref = oracle.grouped_pathwise_greeks(req, 0)
    for g in greeks:
        err = rel_err(res.greeks[g], ref[g.value])
        print(f"{g.value}: max rel err vs oracle pathwise {err:.2e}")
        assert err < 1e-9


def test_origin_exercise_case(hip_f64):
    import finoptions_hip as fh

    schedule = fh.EventSchedule(times=(0.0, 0.5, 1.0),
                                dividends=(fh.CashDividend(node=1, amount=200.0),),
                                exercise_nodes=(0,), expiry_node=2)
    curve = fh.LogDiscountCurve(times=schedule.times, log_discounts=(0.0, 0.0, 0.0))
    method = hip_method(1000, 1000, 7)
    res = fh.price_lsm_call(initial_spot=200.0, strike=100.0, schedule=schedule, curve=curve,
                            dividend_yield_curve=fh.CumulativeDividendYieldCurve.zero(
                                times=schedule.times),
                            volatility=fh.ConstantVolatility(0.30),
                            dividend_policy=fh.DividendPolicy.LIQUIDATOR, executor=hip_f64,
                            method=method)
    assert res.price == 100.0
    assert res.exercise_counts == (1000, 0)
    assert isinstance(res.fit_diagnostics[0], fh.OriginDecisionDiagnostics)
    assert res.fit_diagnostics[0].exercise


def test_european_with_yield_covers_black_scholes(hip_f64):
    import finoptions_hip as fh
    from scipy import stats

    spot, strike, rate, q, vol, t = 100.0, 105.0, 0.05, 0.02, 0.25, 1.0
    schedule = fh.EventSchedule(times=(0.0, t), dividends=(), exercise_nodes=(), expiry_node=1)
    res = fh.price_lsm_call(initial_spot=spot, strike=strike, schedule=schedule,
                            curve=fh.LogDiscountCurve(times=schedule.times,
                                                      log_discounts=(0.0, -rate * t)),
                            dividend_yield_curve=fh.CumulativeDividendYieldCurve.constant(
                                times=schedule.times, annualized_decimal=q),
                            volatility=fh.ConstantVolatility(vol),
                            dividend_policy=fh.DividendPolicy.LIQUIDATOR, executor=hip_f64,
                            method=hip_method(131_072, 131_072, 31))
    d1 = (math.log(spot / strike) + (rate - q + 0.5 * vol * vol) * t) / (vol * math.sqrt(t))
    bs = (spot * math.exp(-q * t) * stats.norm.cdf(d1)
          - strike * math.exp(-rate * t) * stats.norm.cdf(d1 - vol * math.sqrt(t)))
    assert abs(res.price - bs) < 5 * res.standard_error


def test_fail_policy_reports_insufficient_support_like_the_oracle(hip_f64, oracle):
    import finoptions_hip as fh

    cols = c2_columns(strikes=64)
    with pytest.raises(fh.LsmError) as gpu_err:
        _grouped_pair(hip_f64, oracle, cols, base_intervals=50, paths=32_768, seed=11,
                      continue_support=False)
    with pytest.raises(oracle.OracleError) as cpu_err:
        oracle.price_grouped(oracle_grouped(cols, base_intervals=50, training_paths=32_768,
                                            evaluation_paths=32_768, seed=11))
    assert gpu_err.value.code == cpu_err.value.code
    assert gpu_err.value.row == cpu_err.value.row
    assert gpu_err.value.node == cpu_err.value.node
    assert gpu_err.value.index == cpu_err.value.index


def test_condition_limit_failure_is_reported_like_the_oracle(hip_f64, hip_fused, oracle):
    """A condition limit that some fitted rules exceed fails the request at the first failing
    (row, node) in processing order. The fused method accepts each rule provisionally and runs
    the rank/condition gates after the backward pass, so it must report the same failure."""
    import finoptions_hip as fh

    cols = c2_columns(strikes=8)
    req = oracle_grouped(cols, base_intervals=10, training_paths=8192, evaluation_paths=8192,
                         seed=5, insufficient_support=1, condition_limit=140.0)
    with pytest.raises(oracle.OracleError) as cpu_err:
        oracle.price_grouped(req)
    assert cpu_err.value.code == 303
    mesh = fh.UniformExerciseMesh(base_intervals=10, origin_exercise=fh.OriginExercise.EXCLUDED)
    base = hip_method(8192, 8192, 5, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    method = fh.Lsm(training_paths=base.training_paths, evaluation_paths=base.evaluation_paths,
                    training_stream=base.training_stream, evaluation_stream=base.evaluation_stream,
                    insufficient_exercise_support=base.insufficient_exercise_support,
                    condition_limit=140.0)
    for hip in (hip_f64, hip_fused):
        with pytest.raises(fh.LsmError) as gpu_err:
            fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=hip,
                                            method=method)
        e, c = gpu_err.value, cpu_err.value
        assert (e.code, e.row, e.node, e.index) == (c.code, c.row, c.node, c.index)
        assert abs(e.observed - c.observed) <= 1e-10 * c.observed


def _ragged_requests(fh, oracle, contracts, *, seed, path_scale, max_exercise):
    from finoptions_hip.workloads import ragged_batch

    w = ragged_batch(contracts, seed=seed, max_exercise=max_exercise, path_scale=path_scale)
    rows = len(w["initial_spots"])
    dividends = fh.CashDividendBatch(
        schedules=fh.RaggedCashDividends(offsets=w["cash_offsets"], times=w["cash_times"],
                                         amounts=w["cash_amounts"]),
        policy=fh.DividendPolicy.LIQUIDATOR, maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    mesh = fh.ExplicitExerciseMesh(
        schedules=fh.RaggedExerciseTimes(offsets=w["exercise_offsets"], times=w["exercise_times"]),
        origin_exercise=fh.OriginExercise.EXCLUDED)
    counts = fh.RaggedPathCounts(training=w["paths"], evaluation=w["paths"])
    method = hip_method(int(w["paths"].max()), int(w["paths"].max()), 2024,
                        support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    req = oracle.GroupedRequest(
        group_offsets=np.arange(rows + 1, dtype=np.uint32), initial_spots=w["initial_spots"],
        strikes=w["strikes"], maturities=w["maturities"], rates=w["rates"],
        volatilities=w["volatilities"], cash_offsets=w["cash_offsets"],
        cash_times=w["cash_times"], cash_amounts=w["cash_amounts"], cash_policy=0,
        training_paths=int(w["paths"].max()), evaluation_paths=int(w["paths"].max()),
        training_path_counts=w["paths"], evaluation_path_counts=w["paths"],
        training_seed=2024, evaluation_seed=2024, training_domain_start=1,
        evaluation_domain_start=2 + rows, mesh_kind=2, exercise_offsets=w["exercise_offsets"],
        exercise_times=w["exercise_times"], origin_exercise=0, insufficient_support=1)
    method = fh.Lsm(training_paths=method.training_paths, evaluation_paths=method.evaluation_paths,
                    training_stream=fh.RandomStream(seed=2024, domain=1),
                    evaluation_stream=fh.RandomStream(seed=2024, domain=2 + rows),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return w, dividends, mesh, counts, method, req


def test_ragged_batch_matches_the_oracle(hip_f64, hip_fused, oracle):
    import finoptions_hip as fh

    w, dividends, mesh, counts, method, req = _ragged_requests(fh, oracle, 40, seed=2024,
                                                               path_scale=1 / 32,
                                                               max_exercise=40)
    kwargs = dict(initial_spots=w["initial_spots"], strikes=w["strikes"],
                  maturities=w["maturities"], rates=w["rates"], volatilities=w["volatilities"],
                  dividends=dividends, mesh=mesh, method=method, path_counts=counts)
    res = fh.price_lsm_call_batch(executor=hip_f64, **kwargs)
    ref = oracle.price_grouped(req)
    err = rel_err(res.prices, ref["prices"])
    print(f"ragged batch (40 contracts): FP64 max price rel err {err:.2e}")
    assert err < PRICE_TOL
    assert np.array_equal(res.expiry_paths, ref["expiry_paths"])
    assert np.array_equal(res.early_exercise_paths + res.expiry_paths, w["paths"])
    bulk = fh.price_lsm_call_batch(executor=hip_fused, **kwargs)
    bias = np.abs(bulk.prices - ref["prices"]) / ref["standard_errors"]
    print(f"ragged batch fused FP64: max |bias|/SE {bias.max():.2e}, "
          f"rel err {rel_err(bulk.prices, ref['prices']):.2e}")
    assert bias.max() < 0.01


def test_large_grouped_frame_slice_matches_the_oracle(hip_f64, hip_fused, oracle):
    import finoptions_hip as fh
    from finoptions_hip.workloads import large_grouped_frame

    w = large_grouped_frame(groups=6, strikes_per_group=24)
    cols = {k: w[k] for k in ("group_offsets", "initial_spots", "strikes", "maturities", "rates",
                              "volatilities", "cash_offsets", "cash_times", "cash_amounts")}
    # groups take consecutive stream domains: training 1..6, evaluation 7..12
    res_pair = _grouped_pair(hip_f64, oracle, cols, base_intervals=w["intervals"], paths=4096,
                             seed=5, eval_domain=7)
    res, ref = res_pair[0], res_pair[1]
    err = rel_err(res.prices, ref["prices"])
    print(f"grouped frame slice (6 x 24): FP64 max price rel err {err:.2e}")
    assert err < PRICE_TOL
    assert np.array_equal(res.expiry_paths, ref["expiry_paths"])
    bulk = _grouped_pair(hip_fused, oracle, cols, base_intervals=w["intervals"], paths=4096,
                         seed=5, eval_domain=7)[0]
    bias = np.abs(bulk.prices - ref["prices"]) / ref["standard_errors"]
    print(f"grouped frame slice fused FP64: max |bias|/SE {bias.max():.2e}")
    assert bias.max() < 0.01


@pytest.mark.parametrize("shape", ["c2-slice", "cash-survivor", "escrowed", "ragged", "tie-window"])
def test_fused_cashflow_encodings_are_bitwise_equal(hip_fused, shape):
    """The fused backward induction keeps each (strike, path) cashflow either as a binary64 value
    column or as a stopping byte from which it is rebuilt with the same operations: every result
    column, Greeks and tie counts included, is bitwise the same under both encodings."""
    import finoptions_hip as fh

    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    if shape == "ragged":
        from finoptions_hip.workloads import ragged_batch, ragged_batch_request

        w = ragged_batch(24, seed=2024, max_exercise=60, path_scale=1 / 16)

        def run():
            return fh.price_lsm_call_batch(executor=hip_fused, greeks=greeks,
                                           **ragged_batch_request(w))
    else:
        cols = c2_columns(strikes=24)
        if shape in ("cash-survivor", "escrowed"):
            del cols["yields"]
            cols.update(cash_offsets=np.array([0, 3], dtype=np.int64),
                        cash_times=np.array([0.25, 0.5, 0.75]),
                        cash_amounts=np.array([5.0, 8.0, 10.0]))
            if shape == "cash-survivor":
                cols["cash_policy"] = 1
            else:
                cols["escrowed"] = 1
        mesh = fh.UniformExerciseMesh(base_intervals=30, origin_exercise=fh.OriginExercise.EXCLUDED)
        method = hip_method(20_000, 20_000, 29, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)

        def run():
            return fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh,
                                                   executor=hip_fused, method=method, greeks=greeks)
    if shape == "tie-window":
        hip_fused.set_near_tie_window(2.0 ** -20)
    try:
        outs = {}
        for enc in ("value_columns", "stopping_rows"):
            hip_fused.tune(fit_cashflows=enc)
            outs[enc] = run()
    finally:
        hip_fused.tune()
        hip_fused.set_near_tie_window(0.0)
    a, b = outs["value_columns"], outs["stopping_rows"]
    for name in ("prices", "sample_standard_deviations", "standard_errors",
                 "maximum_regression_condition_numbers", "early_exercise_paths", "expiry_paths",
                 "regression_nodes", "minimum_in_the_money_training_paths",
                 "near_tie_training_decisions", "near_tie_evaluation_decisions"):
        x, y = np.asarray(getattr(a, name)), np.asarray(getattr(b, name))
        if x.dtype == np.float64:
            x, y = x.view(np.uint64), y.view(np.uint64)
        assert np.array_equal(x, y), (shape, name)
    for g in greeks:
        assert np.array_equal(a.greeks[g].view(np.uint64), b.greeks[g].view(np.uint64)), (shape, g)
    if shape == "tie-window":
        assert int(a.near_tie_training_decisions.sum()) > 0
