"""PyTorch autograd surface: grouped LSM prices differentiable in S0, sigma, r and q.

``lsm_grouped_prices(spots, vols, rates, yields, ...)`` returns per-strike-row prices as a tensor.
Its backward is the vector-Jacobian product of the pathwise sensitivities under each strike's
frozen stopping policy: d price / d S0, d sigma, d r and d q. The evaluation kernel computes them
from paths regenerated from their Philox counters, in the same pass as the prices. Gradients
reach the per-group inputs by summing each group's strike rows.

The pricing itself runs in the HIP library on the executor's own stream; torch tensors only carry
inputs and outputs (host round trip of a few numbers per group).
"""

from __future__ import annotations

import numpy as np
import torch

from .batch_lsm import (
    ContinuousDividendYieldBatch,
    ExerciseMesh,
    LsmStrikeGroups,
    price_lsm_call_grouped_batch,
)
from .lsm import Greek, Lsm
from .simulation import HipExecutor

_ALL = (Greek.DELTA, Greek.VEGA, Greek.RHO, Greek.DIVIDEND_YIELD)


class _GroupedLsmPrice(torch.autograd.Function):
    @staticmethod
    def forward(ctx, spots, vols, rates, yields, group_offsets, strikes, maturities, mesh,
                executor, method):
        host = [t.detach().to("cpu", torch.float64).contiguous().numpy()
                for t in (spots, vols, rates, yields)]
        groups = LsmStrikeGroups(
            group_offsets=group_offsets, initial_spots=host[0], strikes=strikes,
            maturities=maturities, rates=host[2], volatilities=host[1],
            dividends=ContinuousDividendYieldBatch(annualized_decimals=host[3]))
        result = price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=executor,
                                              method=method, greeks=_ALL)
        rows_per_group = np.diff(group_offsets.astype(np.int64))
        ctx.rows_per_group = torch.as_tensor(rows_per_group)
        sens = np.stack([result.greeks[g] for g in _ALL])        # [4, rows]
        ctx.save_for_backward(torch.as_tensor(sens, device=spots.device))
        ctx.result = result
        return torch.as_tensor(np.array(result.prices), dtype=torch.float64, device=spots.device)

    @staticmethod
    def backward(ctx, grad_prices):
        (sens,) = ctx.saved_tensors
        weighted = sens * grad_prices.to(sens.dtype).unsqueeze(0)    # [4, rows]
        index = torch.repeat_interleave(torch.arange(ctx.rows_per_group.numel()),
                                        ctx.rows_per_group).to(sens.device)
        groups = ctx.rows_per_group.numel()
        per_group = torch.zeros(4, groups, dtype=sens.dtype, device=sens.device)
        per_group.index_add_(1, index, weighted)
        d_spot, d_vol, d_rate, d_yield = per_group
        return (d_spot, d_vol, d_rate, d_yield, None, None, None, None, None, None)


def lsm_grouped_prices(spots: torch.Tensor, vols: torch.Tensor, rates: torch.Tensor,
                       yields: torch.Tensor, *, group_offsets: np.ndarray, strikes: np.ndarray,
                       maturities: np.ndarray, mesh: ExerciseMesh, executor: HipExecutor,
                       method: Lsm) -> torch.Tensor:
    """Differentiable grouped LSM call prices (continuous dividend yield per group)."""
    return _GroupedLsmPrice.apply(spots, vols, rates, yields, group_offsets, strikes, maturities,
                                  mesh, executor, method)


__all__ = ["lsm_grouped_prices"]
