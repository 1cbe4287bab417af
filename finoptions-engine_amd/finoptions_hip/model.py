"""Financial inputs of a single scheduled call: dividends, volatility, event grid and curves.

The classes and fields are those of the reference's python/finoptions/model.py, so requests port
unchanged. Runtime types are checked through the rule tables of ``_rules``; the numerical
contract of every grid, curve and amount (finiteness, ordering, canonical zeros, representable
intervals) is decided by libfohip's ``fohip_validate_*`` entry points, the same host code the
pricing calls run, so a value accepted here is accepted there.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum

import numpy as np

from . import _rules as R
from ._lib import FohipError, InputError, lib, ptr, raise_for


def _native_check(entry: str, *args) -> None:
    """Run one fohip_validate_* entry point; a failure becomes its InputError."""
    err = FohipError()
    raise_for(getattr(lib(), entry)(*args, C.byref(err)), err, lsm=False)


def _doubles(values) -> np.ndarray:
    # the C side never dereferences a zero-length array, but ctypes needs a valid pointer
    return np.ascontiguousarray(list(values) or [0.0], dtype=np.float64)


def _words(values) -> np.ndarray:
    return np.ascontiguousarray(list(values) or [0], dtype=np.uint32)


class DividendPolicy(str, Enum):
    """What a cash dividend does to a spot it exceeds: LIQUIDATOR pays what is there and leaves
    zero, SURVIVOR skips the payment and leaves the spot unchanged."""

    LIQUIDATOR = "liquidator"
    SURVIVOR = "survivor"


@dataclass(frozen=True, slots=True)
class CashDividend:
    """Cash amount paid at one node of an event grid (node index, amount >= 0)."""

    node: int
    amount: float

    def __post_init__(self) -> None:
        R.enforce([("cash-dividend node", self.node, R.PLAIN_U32),
                   ("cash-dividend amount", self.amount,
                    (R.exact_type(float, "{field} must be a plain float; observed {kind}"),))])
        _native_check("fohip_validate_cash_dividend_amount", self.amount)


@dataclass(frozen=True, slots=True)
class ConstantVolatility:
    """Annualized volatility of the GBM, one value for the whole grid (finite, > 0)."""

    annualized_decimal: float

    def __post_init__(self) -> None:
        R.enforce([("constant annualized volatility", self.annualized_decimal,
                    (R.exact_type(float, "{field} must be a plain decimal float; observed {kind}"),))])
        _native_check("fohip_validate_constant_volatility", self.annualized_decimal)


@dataclass(frozen=True, slots=True)
class EventSchedule:
    """Grid times, the cash dividends on it, the exercise nodes and the expiry node. An exercise
    node sees the state before a dividend paid at the same node; expiry sees it after."""

    times: tuple[float, ...]
    dividends: tuple[CashDividend, ...]
    exercise_nodes: tuple[int, ...]
    expiry_node: int

    def __post_init__(self) -> None:
        R.enforce([
            ("event times", self.times, R.FLOAT_TUPLE),
            ("dividends", self.dividends,
             (R.Check(lambda d: type(d) is tuple and all(type(x) is CashDividend for x in d),
                      "dividends must be an immutable tuple[CashDividend, ...]"),)),
            ("exercise nodes", self.exercise_nodes, R.U32_TUPLE),
            ("expiry node", self.expiry_node, R.PLAIN_U32),
        ])
        t = _doubles(self.times)
        dn = _words(d.node for d in self.dividends)
        da = _doubles(d.amount for d in self.dividends)
        en = _words(self.exercise_nodes)
        _native_check("fohip_validate_event_schedule",
                      ptr(t if self.times else None, C.c_double), len(self.times),
                      ptr(dn, C.c_uint32), ptr(da, C.c_double), len(self.dividends),
                      ptr(en, C.c_uint32), len(self.exercise_nodes), self.expiry_node)


def _curve_rows(times_field: str, times, values_field: str, values, what: str) -> list:
    """Rows of a curve: its times, its values, then one value per node."""
    return [(times_field, times, R.FLOAT_TUPLE), (values_field, values, R.FLOAT_TUPLE),
            (values_field, (times, values),
             (R.Lazy(lambda tv: len(tv[0]) == len(tv[1]),
                     lambda tv: f"{what} has {len(tv[1])} values for {len(tv[0])} nodes"),))]


@dataclass(frozen=True, slots=True)
class LogDiscountCurve:
    """Natural logarithms of the discount factors P(0, t) at the grid times (0 at t = 0)."""

    times: tuple[float, ...]
    log_discounts: tuple[float, ...]

    def __post_init__(self) -> None:
        R.enforce(_curve_rows("curve times", self.times, "log discounts", self.log_discounts,
                              "log discount curve"))
        _native_check("fohip_validate_log_discount_curve", ptr(_doubles(self.times), C.c_double),
                      ptr(_doubles(self.log_discounts), C.c_double), len(self.times))


@dataclass(frozen=True, slots=True)
class CumulativeDividendYieldCurve:
    """Integrated proportional dividend yield Q(t) at the grid times (0 at t = 0, nondecreasing):
    a spot grows by exp(-dlogD - dQ) in expectation between two nodes."""

    times: tuple[float, ...]
    cumulative_yields: tuple[float, ...]

    def __post_init__(self) -> None:
        R.enforce(_curve_rows("dividend-yield times", self.times, "cumulative dividend yields",
                              self.cumulative_yields, "dividend-yield curve"))
        _native_check("fohip_validate_cumulative_yield_curve", ptr(_doubles(self.times), C.c_double),
                      ptr(_doubles(self.cumulative_yields), C.c_double), len(self.times))

    @classmethod
    def constant(cls, *, times: tuple[float, ...], annualized_decimal: float) -> "CumulativeDividendYieldCurve":
        """Q(t) = q t on ``times`` for a constant annualized yield q >= 0."""
        R.enforce([("constant dividend yield", annualized_decimal,
                    (R.Check(R.canonical_nonnegative,
                             "{field} must be a finite canonical nonnegative plain float; "
                             "observed {value!r}"),)),
                   ("dividend-yield times", times, R.FLOAT_TUPLE)])
        return cls(times=times, cumulative_yields=tuple(annualized_decimal * t for t in times))

    @classmethod
    def zero(cls, *, times: tuple[float, ...]) -> "CumulativeDividendYieldCurve":
        """Q(t) = 0 on ``times``: no proportional dividend yield."""
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
