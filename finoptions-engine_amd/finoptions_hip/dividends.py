"""Arrow-style ragged cash-dividend schedules consumed by batched LSM pricing.

Mirrors ``RaggedCashDividends`` and ``MaturityBoundary`` of python/finoptions/dividends.py
(validation rules of crates/finoptions-batch/src/view.rs:48-254).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum
from typing import cast

import numpy as np
import numpy.typing as npt

from ._lib import FohipError, InputError, lib, ptr, raise_for

Float64Vector = npt.NDArray[np.float64]
Int64Vector = npt.NDArray[np.int64]


def validate_vector(value: object, *, field: str, dtype: str) -> None:
    """Require a 1-D C-contiguous NumPy array of an exact dtype (no implicit copies)."""
    if type(value) is not np.ndarray:
        raise InputError(f"{field} must be a plain numpy.ndarray")
    array = cast("npt.NDArray[np.generic]", value)
    expected = np.dtype(dtype)
    if array.dtype != expected:
        raise InputError(f"{field} must have dtype {expected}; observed {array.dtype}")
    if array.ndim != 1:
        raise InputError(f"{field} must be one-dimensional; observed shape {array.shape}")
    if not array.flags.c_contiguous:
        raise InputError(f"{field} must be C-contiguous; implicit copies are forbidden")


class MaturityBoundary(str, Enum):
    """Whether a cash dividend dated exactly at maturity is admitted (AT_OR_BEFORE; paid before the
    expiry payoff) or rejected (STRICTLY_BEFORE)."""

    STRICTLY_BEFORE = "strictly_before"
    AT_OR_BEFORE = "at_or_before"


@dataclass(frozen=True, slots=True)
class RaggedCashDividends:
    """Per-row cash dividends in Arrow list layout: int64 ``offsets`` (rows + 1 entries, possibly
    starting above 0) into float64 ``times`` (positive, nondecreasing within a row) and
    ``amounts`` (finite, >= 0). The arrays are borrowed, not copied."""

    offsets: Int64Vector
    times: Float64Vector
    amounts: Float64Vector

    def __post_init__(self) -> None:
        validate_vector(self.offsets, field="offsets", dtype="int64")
        validate_vector(self.times, field="times", dtype="float64")
        validate_vector(self.amounts, field="amounts", dtype="float64")
        if self.times.size != self.amounts.size:
            raise InputError(
                f"cash-dividend times ({self.times.size}) and amounts ({self.amounts.size}) differ")
        err = FohipError()
        code = lib().fohip_validate_ragged_offsets(ptr(self.offsets, C.c_int64), self.offsets.size,
                                                   self.times.size, C.byref(err))
        raise_for(code, err)
        code = lib().fohip_validate_ragged_cash_dividends(
            ptr(self.offsets, C.c_int64), self.offsets.size, ptr(self.times, C.c_double),
            ptr(self.amounts, C.c_double), C.byref(err))
        raise_for(code, err)

    @property
    def rows(self) -> int:
        return int(self.offsets.size - 1)

    @property
    def events(self) -> int:
        return int(self.times.size)


__all__ = ["MaturityBoundary", "RaggedCashDividends", "validate_vector"]
