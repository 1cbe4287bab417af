"""Pathwise Greeks (frozen policy) against common-random-number central finite differences and
through the torch.autograd surface."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method

pytestmark = pytest.mark.gpu

REL_BUMP = 1e-4


def _price(fh, executor, cols, paths, greeks=(), intervals=10):
    mesh = fh.UniformExerciseMesh(base_intervals=intervals,
                                  origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(paths, paths, 21, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_pathwise_greeks_agree_with_crn_finite_differences(precision, hip_f64, hip_f32):
    import finoptions_hip as fh

    executor = hip_f64 if precision == "f64" else hip_f32
    paths = 65_536
    base = c2_columns(strikes=8)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)
    res = _price(fh, executor, base, paths, greeks)
    fields = {fh.Greek.DELTA: "initial_spots", fh.Greek.VEGA: "volatilities",
              fh.Greek.RHO: "rates", fh.Greek.DIVIDEND_YIELD: "yields"}
    for greek, name in fields.items():
        h = REL_BUMP * float(base[name][0])
        up = {**base, name: base[name] + h}
        dn = {**base, name: base[name] - h}
        fd = (_price(fh, executor, up, paths).prices - _price(fh, executor, dn, paths).prices) / (2 * h)
        pw = res.greeks[greek]
        se = res.greek_standard_errors[greek]
        z = np.abs(fd - pw) / np.maximum(se, 1e-12)
        print(f"{precision} {greek.value}: pathwise {pw[:3]} fd {fd[:3]} max |fd-pw|/SE {z.max():.2f}")
        assert np.all(z < 6.0), (greek, fd, pw, se)


def test_autograd_backward_is_the_pathwise_vjp(hip_f64):
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
    method = hip_method(16_384, 16_384, 4, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    leaves = [torch.tensor(cols[k], dtype=torch.float64, requires_grad=True)
              for k in ("initial_spots", "volatilities", "rates", "yields")]
    prices = lsm_grouped_prices(*leaves, group_offsets=cols["group_offsets"],
                                strikes=cols["strikes"], maturities=cols["maturities"], mesh=mesh,
                                executor=hip_f64, method=method)
    weights = torch.linspace(0.5, 1.5, prices.numel(), dtype=torch.float64)
    (prices * weights).sum().backward()
    direct = fh.price_lsm_call_grouped_batch(
        groups=hip_groups(cols), mesh=mesh, executor=hip_f64, method=method,
        greeks=(fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD))
    np.testing.assert_array_equal(prices.detach().numpy(), direct.prices)
    w = weights.numpy()
    for leaf, greek in zip(leaves, (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO,
                                    fh.Greek.DIVIDEND_YIELD)):
        g = direct.greeks[greek] * w
        expected = np.array([g[0:3].sum(), g[3:8].sum()])
        np.testing.assert_allclose(leaf.grad.numpy(), expected, rtol=1e-12)
