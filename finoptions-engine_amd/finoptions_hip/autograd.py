"""PyTorch autograd surface: grouped LSM prices differentiable in S0, sigma, r and q.

``lsm_grouped_prices(spots, vols, rates, yields, ...)`` returns the per-strike-row prices of a
grouped request as a float64 tensor. Its backward is the vector-Jacobian product of the pathwise
sensitivities under each strike's frozen stopping policy (d price / d S0, d sigma, d r, d q), which
the evaluation kernel computes from the regenerated evaluation paths in the pricing pass itself.
Only the sensitivities of inputs that require grad are requested from the device; a forward with
no such input prices without Greeks. Gradients reach the per-group inputs by summing each group's
strike rows.

The dividends are either a per-group continuous yield (the ``yields`` tensor, differentiable) or,
with ``yields=None``, cash dividends given as ``dividends=CashDividendBatch(...)`` or
``EscrowedCashDividendBatch(...)``, which carry no gradient: those groups differentiate in S0,
sigma and r.

The pricing runs in the HIP library on the executor's stream; tensors only carry inputs and
outputs (a host round trip of a few numbers per group).
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from ._lib import InputError
from .batch_lsm import (
    CashDividendBatch,
    ContinuousDividendYieldBatch,
    EscrowedCashDividendBatch,
    ExerciseMesh,
    LsmStrikeGroups,
    RaggedPathCounts,
    price_lsm_call_grouped_batch,
)
from .lsm import Greek, Lsm
from .simulation import HipExecutor

# the differentiable inputs in argument order and the sensitivity each one needs
_INPUT_GREEKS = (Greek.DELTA, Greek.VEGA, Greek.RHO, Greek.DIVIDEND_YIELD)


@dataclass(frozen=True)
class _Request:
    group_offsets: np.ndarray
    strikes: np.ndarray
    maturities: np.ndarray
    mesh: ExerciseMesh
    executor: HipExecutor
    method: Lsm
    dividends: CashDividendBatch | EscrowedCashDividendBatch | None
    path_counts: RaggedPathCounts | None


def _host(t: torch.Tensor) -> np.ndarray:
    return np.ascontiguousarray(t.detach().to("cpu", torch.float64).numpy())


class _GroupedLsmPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spots, vols, rates, yields, request: _Request):
        wanted = tuple(g for g, need in zip(_INPUT_GREEKS, ctx.needs_input_grad[:4]) if need)
        dividends = (ContinuousDividendYieldBatch(annualized_decimals=_host(yields))
                     if yields is not None else request.dividends)
        groups = LsmStrikeGroups(group_offsets=request.group_offsets, initial_spots=_host(spots),
                                 strikes=request.strikes, maturities=request.maturities,
                                 rates=_host(rates), volatilities=_host(vols), dividends=dividends)
        result = price_lsm_call_grouped_batch(groups=groups, mesh=request.mesh,
                                              executor=request.executor, method=request.method,
                                              greeks=wanted, path_counts=request.path_counts)
        rows_per_group = np.diff(request.group_offsets.astype(np.int64))
        ctx.rows_per_group = torch.as_tensor(rows_per_group)
        ctx.wanted = wanted
        sens = [torch.as_tensor(result.greeks[g], device=spots.device) for g in wanted]
        ctx.save_for_backward(*sens)
        ctx.result = result
        return torch.as_tensor(result.prices, dtype=torch.float64, device=spots.device)

    @staticmethod
    def backward(ctx, grad_prices):
        sens = dict(zip(ctx.wanted, ctx.saved_tensors))
        groups = ctx.rows_per_group.numel()
        index = None
        grads = []
        for g in _INPUT_GREEKS:
            if g not in sens:
                grads.append(None)
                continue
            if index is None:
                index = torch.repeat_interleave(torch.arange(groups), ctx.rows_per_group).to(grad_prices.device)
            weighted = sens[g] * grad_prices.to(sens[g].dtype)
            per_group = torch.zeros(groups, dtype=weighted.dtype, device=weighted.device)
            per_group.index_add_(0, index, weighted)
            grads.append(per_group)
        return (*grads, None)


def lsm_grouped_prices(spots: torch.Tensor, vols: torch.Tensor, rates: torch.Tensor,
                       yields: torch.Tensor | None = None, *, group_offsets: np.ndarray,
                       strikes: np.ndarray, maturities: np.ndarray, mesh: ExerciseMesh,
                       executor: HipExecutor, method: Lsm,
                       dividends: CashDividendBatch | EscrowedCashDividendBatch | None = None,
                       path_counts: RaggedPathCounts | None = None) -> torch.Tensor:
    """Differentiable grouped LSM call prices, one per strike row.

    ``spots``, ``vols``, ``rates`` (and ``yields``) are per-group float tensors. Give either
    ``yields`` (continuous dividend yield, differentiable) or ``dividends`` (cash dividends, no
    gradient), not both.
    """
    if (yields is None) == (dividends is None):
        raise InputError("give either yields (continuous dividend yield) or dividends (cash dividends)")
    if dividends is not None and type(dividends) not in (CashDividendBatch, EscrowedCashDividendBatch):
        raise InputError("dividends must be CashDividendBatch or EscrowedCashDividendBatch")
    for name, t in (("spots", spots), ("vols", vols), ("rates", rates), ("yields", yields)):
        if t is not None and (not isinstance(t, torch.Tensor) or t.ndim != 1):
            raise InputError(f"{name} must be a one-dimensional torch.Tensor")
    request = _Request(group_offsets=group_offsets, strikes=strikes, maturities=maturities, mesh=mesh,
                       executor=executor, method=method, dividends=dividends, path_counts=path_counts)
    return _GroupedLsmPrice.apply(spots, vols, rates, yields, request)


__all__ = ["lsm_grouped_prices"]
