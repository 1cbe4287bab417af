"""Adjoint Greeks on the headline contracts: the single scheduled call (S0 = K = 100, times
(0, 0.5, 1), cash 2.0 at node 1, 65 536 + 65 536 paths, streams (7, 1)/(7, 2)) and all 64 strikes
of the grouped-strike state over 50 intervals, against the oracle's frozen-policy pathwise
estimator; and the pathwise estimator against bump-and-reprice (the plain FP64 method re-fitted
at every bump, common random numbers), reported statistically over independent replications.

Bars: FP64 pathwise Greeks within 1e-9 relative of the oracle; the fused method (binary32
tangents) within 0.1 of the FP64 Greek's standard error.
"""

from __future__ import annotations

import math

import numpy as np
import pytest

from conftest import c1_inputs, c1_plan, c2_columns, hip_groups, oracle_grouped, rel_err

pytestmark = pytest.mark.gpu

GREEK_NAMES = ("DELTA", "VEGA", "RHO", "DIVIDEND_YIELD")


def _c1_method(fh, paths=65_536, seeds=(7, 7)):
    return fh.Lsm(training_paths=paths, evaluation_paths=paths,
                  training_stream=fh.RandomStream(seed=seeds[0], domain=1),
                  evaluation_stream=fh.RandomStream(seed=seeds[1], domain=2),
                  insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL)


def test_c1_contract_greeks_match_the_oracle(hip_f64, hip_fused, oracle):
    import finoptions_hip as fh

    greeks = tuple(getattr(fh.Greek, g) for g in GREEK_NAMES)
    res = fh.price_lsm_call(**c1_inputs(), executor=hip_f64, method=_c1_method(fh), greeks=greeks)
    ref = oracle.call_pathwise_greeks(c1_plan(), 100.0, 100.0, 65_536, 65_536, 7, 1, 7, 2)
    for g in greeks:
        err = rel_err(res.greek(g), ref[g.value])
        print(f"C1 {g.value}: HIP {res.greek(g):.12f} oracle {ref[g.value]:.12f} rel {err:.1e} "
              f"(SE {res.greek_standard_error(g):.3e})")
        assert err < 1e-9, g
    fused = fh.price_lsm_call(**c1_inputs(), executor=hip_fused, method=_c1_method(fh),
                              greeks=greeks)
    for g in greeks:
        z = abs(fused.greek(g) - ref[g.value]) / res.greek_standard_error(g)
        print(f"C1 fused {g.value}: |bias|/SE {z:.2e}")
        assert z < 0.1, g
    assert abs(fused.price - res.price) / res.standard_error < 0.1


def _c2(fh, executor, paths, greeks=(), *, spot=100.0, vol=0.25, rate=0.03, q=0.02,
        seeds=(11, 11)):
    cols = c2_columns(strikes=64, spot=spot)
    cols["volatilities"] = np.array([vol])
    cols["rates"] = np.array([rate])
    cols["yields"] = np.array([q])
    mesh = fh.UniformExerciseMesh(base_intervals=50, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = fh.Lsm(training_paths=paths, evaluation_paths=paths,
                    training_stream=fh.RandomStream(seed=seeds[0], domain=1),
                    evaluation_stream=fh.RandomStream(seed=seeds[1], domain=2),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


def test_c2_greeks_of_all_64_strikes_match_the_oracle(hip_f64, hip_fused, oracle):
    import finoptions_hip as fh

    paths = 32_768
    greeks = tuple(getattr(fh.Greek, g) for g in GREEK_NAMES)
    res = _c2(fh, hip_f64, paths, greeks)
    req = oracle_grouped(c2_columns(strikes=64), base_intervals=50, training_paths=paths,
                         evaluation_paths=paths, seed=11, insufficient_support=1)
    ref = oracle.grouped_pathwise_greeks(req, 0)
    fused = _c2(fh, hip_fused, paths, greeks)
    for g in greeks:
        err = rel_err(res.greeks[g], ref[g.value])
        z = np.abs(fused.greeks[g] - ref[g.value]) / res.greek_standard_errors[g]
        print(f"C2 64 x 50 {g.value}: FP64 rel err {err:.2e}; fused max |bias|/SE {z.max():.2e}")
        assert err < 1e-9, g
        assert z.max() < 0.1, g


def _replicated_fd(price_fn, base_value, rel_bump, replications):
    """Per replication r (independent training and evaluation seeds): the pathwise Greek g_r and
    the re-fitted central difference d_r = (P(x + h) - P(x - h)) / 2h under common random numbers.
    Returns the mean of g_r - d_r, its standard error over replications, and the mean SE of g_r."""
    diffs, ses = [], []
    h = rel_bump * base_value
    for r in range(replications):
        g, g_se = price_fn(None, r)
        up = price_fn(base_value + h, r)
        down = price_fn(base_value - h, r)
        diffs.append(g - (up - down) / (2 * h))
        ses.append(g_se)
    diffs = np.array(diffs)
    mean = diffs.mean(axis=0)
    se = diffs.std(axis=0, ddof=1) / math.sqrt(replications)
    return mean, se, np.mean(ses, axis=0)


@pytest.mark.parametrize("rel_bump", [1e-2, 1e-4])
def test_c1_pathwise_greeks_against_bump_and_reprice(hip_f64, rel_bump):
    """The frozen-policy pathwise delta and vega against the re-fitted FP64 central difference
    over 16 independent seed pairs. z = mean(pathwise - FD) / SE(mean); at the 1e-2 bump the
    policy re-fit noise is small against the sampling noise and |z| stays small; at 1e-4 the
    difference is dominated by re-fit noise of order (paths x h)^-1/2, which the replication SE
    measures."""
    import finoptions_hip as fh

    inp = c1_inputs()
    out = {}
    for name, greek in (("delta", fh.Greek.DELTA), ("vega", fh.Greek.VEGA)):
        base = 100.0 if name == "delta" else 0.20

        def price(x, r, name=name, greek=greek):
            kw = dict(inp)
            if x is None:
                res = fh.price_lsm_call(**kw, executor=hip_f64, method=_c1_method(fh, seeds=(101 + r, 201 + r)),
                                        greeks=(greek,))
                return res.greek(greek), res.greek_standard_error(greek)
            if name == "delta":
                kw["initial_spot"] = x
            else:
                kw["volatility"] = fh.ConstantVolatility(x)
            return fh.price_lsm_call(**kw, executor=hip_f64,
                                     method=_c1_method(fh, seeds=(101 + r, 201 + r))).price

        mean, se, greek_se = _replicated_fd(price, base, rel_bump, 16)
        z = float(mean / se)
        out[name] = (float(mean), float(se), float(greek_se), z)
        print(f"C1 {name}, bump {rel_bump:g}: mean(pathwise - FD) {float(mean):+.3e} "
              f"+- {float(se):.3e} (z {z:+.2f}); pathwise SE {float(greek_se):.3e}")
    if rel_bump == 1e-2:
        for name, (mean, se, greek_se, z) in out.items():
            # no systematic difference beyond 4 replication SEs, and inside a tenth of the
            # pathwise estimator's own SE where the FD is precise
            assert abs(z) < 4.0 or abs(mean) < 0.1 * greek_se, (name, out[name])


@pytest.mark.parametrize("rel_bump", [1e-2, 1e-4])
def test_c2_pathwise_greeks_against_bump_and_reprice(hip_f64, rel_bump):
    """The same comparison for the 64 grouped strikes at 2^18 + 2^18 paths over 6 replications,
    delta and vega; the statistic is printed per strike and summarized."""
    import finoptions_hip as fh

    paths = 1 << 18
    for name, greek in (("delta", fh.Greek.DELTA), ("vega", fh.Greek.VEGA)):
        base = 100.0 if name == "delta" else 0.25

        def price(x, r, name=name, greek=greek):
            seeds = (301 + r, 401 + r)
            if x is None:
                res = _c2(fh, hip_f64, paths, (greek,), seeds=seeds)
                return res.greeks[greek], res.greek_standard_errors[greek]
            kw = {"spot": x} if name == "delta" else {"vol": x}
            return _c2(fh, hip_f64, paths, seeds=seeds, **kw).prices

        mean, se, greek_se = _replicated_fd(price, base, rel_bump, 6)
        z = mean / se
        print(f"C2 {name}, bump {rel_bump:g}: |z| median {np.median(np.abs(z)):.2f}, "
              f"max {np.max(np.abs(z)):.2f}; |mean diff| / pathwise SE median "
              f"{np.median(np.abs(mean) / greek_se):.3f}, max {np.max(np.abs(mean) / greek_se):.3f}")
        assert np.all(np.isfinite(z))
