"""Immutable financial specifications validated by the native library.

Mirrors python/finoptions/model.py of the reference: the same classes, fields, runtime type checks
and error classes; validation runs in libfohip's host code (the reference runs it in its Rust
core through ``_native.validate_*``).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum
from math import copysign, isfinite
from typing import TypeVar, cast

import numpy as np

from ._lib import FohipError, InputError, lib, ptr, raise_for


def _require_plain_int(value: object, *, field: str) -> int:
    if type(value) is not int:
        msg = f"{field} must be a plain int; observed {type(value).__name__}"
        raise InputError(msg)
    if value < 0 or value > (2**32 - 1):
        msg = f"{field} must be in the unsigned 32-bit node domain; observed {value}"
        raise InputError(msg)
    return value


def _require_float_tuple(values: object, *, field: str) -> tuple[float, ...]:
    if type(values) is not tuple:
        msg = f"{field} must be an immutable tuple[float, ...]"
        raise InputError(msg)
    if any(type(value) is not float for value in values):
        msg = f"every {field} value must be a plain float"
        raise InputError(msg)
    return cast("tuple[float, ...]", values)


def _require_int_tuple(values: object, *, field: str) -> tuple[int, ...]:
    if type(values) is not tuple:
        msg = f"{field} must be an immutable tuple[int, ...]"
        raise InputError(msg)
    return tuple(_require_plain_int(value, field=field) for value in values)


def _f64(values) -> np.ndarray:
    return np.ascontiguousarray(values if len(values) else [0.0], dtype=np.float64)


def _u32(values) -> np.ndarray:
    return np.ascontiguousarray(values if len(values) else [0], dtype=np.uint32)


def _check(code: int, err: FohipError) -> None:
    raise_for(code, err, lsm=False)


class DividendPolicy(str, Enum):
    """Explicit limited-liability treatment of deterministic cash dividends."""

    LIQUIDATOR = "liquidator"
    SURVIVOR = "survivor"


@dataclass(frozen=True, slots=True)
class CashDividend:
    """A validated nonnegative cash amount attached to an exact node ID."""

    node: int
    amount: float

    def __post_init__(self) -> None:
        _require_plain_int(self.node, field="cash-dividend node")
        if type(self.amount) is not float:
            msg = (
                f"cash-dividend amount must be a plain float; observed {type(self.amount).__name__}"
            )
            raise InputError(msg)
        err = FohipError()
        _check(lib().fohip_validate_cash_dividend_amount(self.amount, C.byref(err)), err)


@dataclass(frozen=True, slots=True)
class ConstantVolatility:
    """Strictly positive finite annualized decimal volatility."""

    annualized_decimal: float

    def __post_init__(self) -> None:
        if type(self.annualized_decimal) is not float:
            msg = (
                "constant annualized volatility must be a plain decimal float; "
                f"observed {type(self.annualized_decimal).__name__}"
            )
            raise InputError(msg)
        err = FohipError()
        _check(lib().fohip_validate_constant_volatility(self.annualized_decimal, C.byref(err)), err)


@dataclass(frozen=True, slots=True)
class EventSchedule:
    """Finite event grid with pre-dividend exercise and separate terminal expiry."""

    times: tuple[float, ...]
    dividends: tuple[CashDividend, ...]
    exercise_nodes: tuple[int, ...]
    expiry_node: int

    def __post_init__(self) -> None:
        times = _require_float_tuple(self.times, field="event times")
        if type(self.dividends) is not tuple or any(
            type(dividend) is not CashDividend for dividend in self.dividends
        ):
            msg = "dividends must be an immutable tuple[CashDividend, ...]"
            raise InputError(msg)
        exercise_nodes = _require_int_tuple(self.exercise_nodes, field="exercise nodes")
        expiry_node = _require_plain_int(self.expiry_node, field="expiry node")
        t = np.ascontiguousarray(times, dtype=np.float64) if times else np.zeros(0)
        dn = _u32([d.node for d in self.dividends])
        da = _f64([d.amount for d in self.dividends])
        en = _u32(list(exercise_nodes))
        err = FohipError()
        code = lib().fohip_validate_event_schedule(
            ptr(t if t.size else None, C.c_double), t.size, ptr(dn, C.c_uint32),
            ptr(da, C.c_double), len(self.dividends), ptr(en, C.c_uint32), len(exercise_nodes),
            expiry_node, C.byref(err))
        _check(code, err)


@dataclass(frozen=True, slots=True)
class LogDiscountCurve:
    """Validated ``log(P(0,t))`` values aligned with an exact time grid."""

    times: tuple[float, ...]
    log_discounts: tuple[float, ...]

    def __post_init__(self) -> None:
        times = _require_float_tuple(self.times, field="curve times")
        log_discounts = _require_float_tuple(self.log_discounts, field="log discounts")
        if len(times) != len(log_discounts):
            msg = (f"log discount curve has {len(log_discounts)} values for {len(times)} nodes")
            raise InputError(msg)
        err = FohipError()
        code = lib().fohip_validate_log_discount_curve(
            ptr(_f64(times), C.c_double), ptr(_f64(log_discounts), C.c_double), len(times),
            C.byref(err))
        _check(code, err)


@dataclass(frozen=True, slots=True)
class CumulativeDividendYieldCurve:
    """Validated ``Q(t) = integral_0^t q(u) du`` values on an exact time grid."""

    times: tuple[float, ...]
    cumulative_yields: tuple[float, ...]

    def __post_init__(self) -> None:
        times = _require_float_tuple(self.times, field="dividend-yield times")
        cumulative = _require_float_tuple(self.cumulative_yields,
                                          field="cumulative dividend yields")
        if len(times) != len(cumulative):
            msg = (f"dividend-yield curve has {len(cumulative)} values for {len(times)} nodes")
            raise InputError(msg)
        err = FohipError()
        code = lib().fohip_validate_cumulative_yield_curve(
            ptr(_f64(times), C.c_double), ptr(_f64(cumulative), C.c_double), len(times),
            C.byref(err))
        _check(code, err)

    @classmethod
    def constant(cls, *, times: tuple[float, ...], annualized_decimal: float) -> "CumulativeDividendYieldCurve":
        """Construct the exact cumulative curve for one constant proportional dividend yield."""
        if (
            type(annualized_decimal) is not float
            or not isfinite(annualized_decimal)
            or annualized_decimal < 0.0
            or (annualized_decimal == 0.0 and copysign(1.0, annualized_decimal) < 0.0)
        ):
            msg = (
                "constant dividend yield must be a finite canonical nonnegative plain float; "
                f"observed {annualized_decimal!r}"
            )
            raise InputError(msg)
        checked = _require_float_tuple(times, field="dividend-yield times")
        return cls(times=checked,
                   cumulative_yields=tuple(annualized_decimal * time for time in checked))

    @classmethod
    def zero(cls, *, times: tuple[float, ...]) -> "CumulativeDividendYieldCurve":
        """Construct an explicit zero-continuous-dividend-yield curve."""
        return cls.constant(times=times, annualized_decimal=0.0)


__all__ = [
    "CashDividend",
    "ConstantVolatility",
    "CumulativeDividendYieldCurve",
    "DividendPolicy",
    "EventSchedule",
    "InputError",
    "LogDiscountCurve",
]
