"""Pathwise Greeks on the GPU: fused-method accuracy against the FP64 method, and the torch.autograd
surface. (That the pathwise estimator is the derivative of the stopped payoff is pinned on the
oracle in test_oracle.py; the FP64 GPU Greeks match the oracle's in test_gpu_parity.py.)"""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method, oracle_grouped

pytestmark = pytest.mark.gpu

ALL_GREEKS = ("DELTA", "VEGA", "RHO", "DIVIDEND_YIELD")


def _price(fh, executor, cols, paths, greeks=(), intervals=10, seed=21):
    mesh = fh.UniformExerciseMesh(base_intervals=intervals,
                                  origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(paths, paths, seed, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


def test_fused_greeks_are_within_a_tenth_of_a_standard_error(hip_f64, hip_fused):
    import finoptions_hip as fh

    cols = c2_columns(strikes=16)
    greeks = tuple(getattr(fh.Greek, g) for g in ALL_GREEKS)
    ref = _price(fh, hip_f64, cols, 65_536, greeks, intervals=25)
    fast = _price(fh, hip_fused, cols, 65_536, greeks, intervals=25)
    bias = np.abs(fast.prices - ref.prices) / ref.standard_errors
    print(f"fused price: max |bias|/SE {bias.max():.4f}")
    assert bias.max() < 0.1
    for greek in greeks:
        se = ref.greek_standard_errors[greek]
        z = np.abs(fast.greeks[greek] - ref.greeks[greek]) / se
        print(f"fused {greek.value}: max |bias|/SE {z.max():.4f}")
        assert np.all(se > 0)
        assert z.max() < 0.1, (greek, fast.greeks[greek], ref.greeks[greek], se)


def test_greeks_do_not_change_prices(hip_f64):
    import finoptions_hip as fh

    cols = c2_columns(strikes=8)
    plain = _price(fh, hip_f64, cols, 16_384)
    with_greeks = _price(fh, hip_f64, cols, 16_384, (fh.Greek.DELTA, fh.Greek.VEGA))
    np.testing.assert_array_equal(plain.prices, with_greeks.prices)
    np.testing.assert_array_equal(plain.standard_errors, with_greeks.standard_errors)
    np.testing.assert_array_equal(plain.early_exercise_paths, with_greeks.early_exercise_paths)


def _cash(cols: dict) -> dict:
    """The same state with two liquidator cash dividends in place of the continuous yield."""
    cols = dict(cols)
    cols.pop("yields")
    cols["cash_offsets"] = np.array([0, 2], dtype=np.int64)
    cols["cash_times"] = np.array([0.3, 0.7])
    cols["cash_amounts"] = np.array([2.0, 3.0])
    return cols


@pytest.mark.parametrize("executor", ["hip_f64", "hip_fused"])
def test_delta_and_vega_alone_equal_those_of_all_four_greeks(executor, request):
    """A request for delta and vega alone carries two tangents per path instead of four; the two
    Greeks must be bitwise those of a request for all four: on a single-strike group with cash
    dividends, on one with a continuous yield, on one 6-strike tile, on three tiles over stored
    normals (with a continuous yield and with escrowed cash dividends), and through the
    single-call entry point."""
    import finoptions_hip as fh
    from conftest import c1_inputs

    ex = request.getfixturevalue(executor)
    both = (fh.Greek.DELTA, fh.Greek.VEGA)
    all4 = tuple(getattr(fh.Greek, g) for g in ALL_GREEKS)
    escrowed = dict(_cash(c2_columns(strikes=16)), escrowed=1)
    for cols in (_cash(c2_columns(strikes=1)), c2_columns(strikes=1), c2_columns(strikes=6),
                 c2_columns(strikes=16), escrowed):
        # 16 strikes: three 6-strike tiles over stored evaluation normals
        greeks = all4 if "yields" in cols else (*both, fh.Greek.RHO)
        two = _price(fh, ex, cols, 16_384, both, intervals=12)
        four = _price(fh, ex, cols, 16_384, greeks, intervals=12)
        for g in both:
            np.testing.assert_array_equal(two.greeks[g], four.greeks[g])
            np.testing.assert_array_equal(two.greek_standard_errors[g],
                                          four.greek_standard_errors[g])
        np.testing.assert_array_equal(two.prices, four.prices)
    inp = c1_inputs()
    method = hip_method(8192, 8192, 7, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    two = fh.price_lsm_call(**inp, executor=ex, method=method, greeks=both)
    four = fh.price_lsm_call(**inp, executor=ex, method=method, greeks=(*both, fh.Greek.RHO))
    assert [g for g, _, _ in two.greeks] == list(both)
    for g in both:
        assert two.greek(g) == four.greek(g)
        assert two.greek_standard_error(g) == four.greek_standard_error(g)


def test_unrequested_greeks_are_nan_in_the_c_result(hip_fused, monkeypatch):
    """fohip_price_lsm_call fills only the requested Greeks and writes NaN into the others."""
    import finoptions_hip as fh
    from conftest import c1_inputs
    from finoptions_hip import lsm as lsm_module

    seen = []

    class Recording(lsm_module.FohipCallResult):
        def __init__(self):
            super().__init__()
            seen.append(self)

    monkeypatch.setattr(lsm_module, "FohipCallResult", Recording)
    res = fh.price_lsm_call(**c1_inputs(), executor=hip_fused,
                            method=hip_method(4096, 4096, 7), greeks=(fh.Greek.VEGA,))
    assert [g for g, _, _ in res.greeks] == [fh.Greek.VEGA]
    (raw,) = seen
    assert np.isfinite(raw.greeks[1]) and raw.greeks[1] == res.greek(fh.Greek.VEGA)
    for i in (0, 2, 3):
        assert np.isnan(raw.greeks[i]) and np.isnan(raw.greek_standard_errors[i])


def _oracle_vjp(oracle, req, rows_per_group, weights, names):
    """Per-group sums of weight x the oracle's pathwise sensitivity, group by group."""
    out = {n: [] for n in names}
    start = 0
    for g, n_rows in enumerate(rows_per_group):
        ref = oracle.grouped_pathwise_greeks(req, g)
        for n in names:
            out[n].append(float(np.sum(ref[n] * weights[start:start + n_rows])))
        start += n_rows
    return {n: np.array(v) for n, v in out.items()}


def test_autograd_gradient_is_the_oracle_pathwise_vjp(hip_f64, oracle):
    """Two continuous-yield groups: backward equals the vector-Jacobian product of the oracle's
    pathwise sensitivities (frozen policy) in S0, sigma, r and q."""
    import finoptions_hip as fh
    import torch
    from finoptions_hip.autograd import lsm_grouped_prices

    cols = c2_columns(strikes=8)
    cols["group_offsets"] = np.array([0, 3, 8], dtype=np.uint32)
    cols["initial_spots"] = np.array([100.0, 90.0])
    cols["maturities"] = np.array([1.0, 1.0])
    cols["rates"] = np.array([0.03, 0.02])
    cols["volatilities"] = np.array([0.25, 0.3])
    cols["yields"] = np.array([0.02, 0.01])
    mesh = fh.UniformExerciseMesh(base_intervals=10, origin_exercise=fh.OriginExercise.EXCLUDED)
    # two groups take training domains 1..2 and evaluation domains 3..4
    method = hip_method(16_384, 16_384, 4, eval_domain=3,
                        support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    leaves = [torch.tensor(cols[k], dtype=torch.float64, requires_grad=True)
              for k in ("initial_spots", "volatilities", "rates", "yields")]
    prices = lsm_grouped_prices(*leaves, group_offsets=cols["group_offsets"],
                                strikes=cols["strikes"], maturities=cols["maturities"], mesh=mesh,
                                executor=hip_f64, method=method)
    weights = torch.linspace(0.5, 1.5, prices.numel(), dtype=torch.float64)
    (prices * weights).sum().backward()
    req = oracle_grouped(cols, base_intervals=10, training_paths=16_384, evaluation_paths=16_384,
                         seed=4, eval_domain=3, insufficient_support=1)
    names = ("delta", "vega", "rho", "dividend_yield")
    ref = _oracle_vjp(oracle, req, (3, 5), weights.numpy(), names)
    for leaf, name in zip(leaves, names):
        err = float(np.max(np.abs(leaf.grad.numpy() - ref[name]) / np.abs(ref[name])))
        print(f"autograd d/d{name}: max rel err vs oracle {err:.2e}")
        assert err < 1e-9


def test_autograd_with_cash_dividends_differentiates_spot_vol_and_rate(hip_f64, oracle):
    """Cash-dividend groups (liquidator): gradients in S0, sigma and r equal the oracle's pathwise
    VJP; only the sensitivities of inputs that require grad are computed."""
    import finoptions_hip as fh
    import torch
    from finoptions_hip.autograd import lsm_grouped_prices

    cols = c2_columns(strikes=6)
    cols["group_offsets"] = np.array([0, 2, 6], dtype=np.uint32)
    cols["initial_spots"] = np.array([100.0, 95.0])
    cols["maturities"] = np.array([1.0, 1.0])
    cols["rates"] = np.array([0.03, 0.04])
    cols["volatilities"] = np.array([0.25, 0.35])
    cols.pop("yields")
    cols["cash_offsets"] = np.array([0, 1, 3], dtype=np.int64)
    cols["cash_times"] = np.array([0.5, 0.3, 0.7])
    cols["cash_amounts"] = np.array([2.0, 1.0, 1.5])
    mesh = fh.UniformExerciseMesh(base_intervals=10, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(16_384, 16_384, 6, eval_domain=3,
                        support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    spots = torch.tensor(cols["initial_spots"], dtype=torch.float64, requires_grad=True)
    vols = torch.tensor(cols["volatilities"], dtype=torch.float64, requires_grad=True)
    rates = torch.tensor(cols["rates"], dtype=torch.float64, requires_grad=True)
    dividends = hip_groups(cols).dividends
    prices = lsm_grouped_prices(spots, vols, rates, group_offsets=cols["group_offsets"],
                                strikes=cols["strikes"], maturities=cols["maturities"], mesh=mesh,
                                executor=hip_f64, method=method, dividends=dividends)
    direct = fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=hip_f64,
                                             method=method)
    np.testing.assert_array_equal(prices.detach().numpy(), direct.prices)
    prices.sum().backward()
    req = oracle_grouped(cols, base_intervals=10, training_paths=16_384, evaluation_paths=16_384,
                         seed=6, eval_domain=3, insufficient_support=1)
    ref = _oracle_vjp(oracle, req, (2, 4), np.ones(6), ("delta", "vega", "rho"))
    for leaf, name in ((spots, "delta"), (vols, "vega"), (rates, "rho")):
        err = float(np.max(np.abs(leaf.grad.numpy() - ref[name]) / np.abs(ref[name])))
        print(f"cash autograd d/d{name}: max rel err vs oracle {err:.2e}")
        assert err < 1e-9
