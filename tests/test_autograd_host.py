"""Host logic of the autograd surface, with the device pricing call replaced by a stub: only the
sensitivities of inputs that require grad are requested, and backward is their vector-Jacobian
product summed per group."""

from __future__ import annotations

import types

import numpy as np
import pytest
import torch


def _stub(calls):
    def price(*, groups, mesh, executor, method, greeks, path_counts):
        calls.append(tuple(greeks))
        rows = groups.strikes.size
        base = np.arange(1, rows + 1, dtype=np.float64)
        sens = {g: base * (k + 2) for k, g in enumerate(_ORDER()) if g in greeks}
        return types.SimpleNamespace(prices=base * 10.0, greeks=sens)
    return price


def _ORDER():
    import finoptions_hip as fh

    return (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)


def _inputs(grad):
    names = ("spots", "vols", "rates", "yields")
    vals = ([100.0, 90.0], [0.2, 0.3], [0.01, 0.02], [0.0, 0.01])
    return [torch.tensor(v, dtype=torch.float64, requires_grad=n in grad) for n, v in zip(names, vals)]


def _call(monkeypatch, leaves, calls, **kw):
    import finoptions_hip as fh
    from finoptions_hip import autograd

    monkeypatch.setattr(autograd, "price_lsm_call_grouped_batch", _stub(calls))
    mesh = fh.UniformExerciseMesh(base_intervals=4, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = fh.Lsm(training_paths=16, evaluation_paths=16, training_stream=fh.RandomStream(seed=1, domain=1),
                    evaluation_stream=fh.RandomStream(seed=1, domain=3),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return autograd.lsm_grouped_prices(*leaves, group_offsets=np.array([0, 2, 5], dtype=np.uint32),
                                       strikes=np.linspace(80.0, 120.0, 5), maturities=np.array([1.0, 1.0]),
                                       mesh=mesh, executor=None, method=method, **kw)


@pytest.mark.parametrize("grad", [("spots",), ("vols", "rates"), ("spots", "vols", "rates", "yields"), ()])
def test_only_inputs_that_require_grad_are_differentiated(monkeypatch, grad):
    calls = []
    leaves = _inputs(grad)
    prices = _call(monkeypatch, leaves, calls)
    order = dict(zip(("spots", "vols", "rates", "yields"), _ORDER()))
    assert calls == [tuple(order[n] for n in ("spots", "vols", "rates", "yields") if n in grad)]
    assert np.array_equal(prices.detach().numpy(), np.arange(1, 6) * 10.0)
    if not grad:
        assert not prices.requires_grad
        return
    w = torch.tensor([1.0, 2.0, 3.0, 4.0, 5.0], dtype=torch.float64)
    (prices * w).sum().backward()
    rows = np.arange(1, 6, dtype=np.float64)
    for k, (name, leaf) in enumerate(zip(("spots", "vols", "rates", "yields"), leaves)):
        if name not in grad:
            assert leaf.grad is None
            continue
        vjp = rows * (k + 2) * w.numpy()
        assert np.allclose(leaf.grad.numpy(), [vjp[:2].sum(), vjp[2:].sum()], rtol=0, atol=0)


def test_cash_dividend_groups_differentiate_in_spot_vol_and_rate(monkeypatch):
    import finoptions_hip as fh

    calls = []
    spots, vols, rates, _ = _inputs(("spots", "vols", "rates"))
    cash = fh.CashDividendBatch(
        schedules=fh.RaggedCashDividends(offsets=np.array([0, 1, 1], dtype=np.int64), times=np.array([0.5]),
                                         amounts=np.array([1.0])),
        policy=fh.DividendPolicy.LIQUIDATOR, maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    prices = _call(monkeypatch, [spots, vols, rates, None], calls, dividends=cash)
    assert calls == [(fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO)]
    prices.sum().backward()
    assert spots.grad is not None and rates.grad is not None
    with pytest.raises(fh.InputError, match="either yields"):
        _call(monkeypatch, _inputs(()), calls, dividends=cash)
    with pytest.raises(fh.InputError, match="either yields"):
        _call(monkeypatch, [spots, vols, rates, None], calls)
