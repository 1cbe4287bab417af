"""Independent-policy least-squares Monte Carlo for explicitly scheduled calls, on HIP devices.

Mirrors python/finoptions/lsm.py of the reference: ``Lsm``, the diagnostic classes, ``LsmResult``
and ``price_lsm_call`` keep their names, fields, defaults and checks. The executor is an explicit
``HipExecutor``; ``greeks`` additionally requests pathwise sensitivities under the frozen policy.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum
from math import isfinite
from typing import Final, Union

import numpy as np

from ._lib import (
    FohipCallRequest,
    FohipCallResult,
    FohipError,
    FohipNodeDiagnostic,
    InputError,
    LsmError,
    lib,
    ptr,
    raise_for,
)
from .model import (
    ConstantVolatility,
    CumulativeDividendYieldCurve,
    DividendPolicy,
    EventSchedule,
    LogDiscountCurve,
)
from .simulation import HipExecutor, RandomStream

_MAX_PATHS: Final = (2**32) - 1
_MAX_POLYNOMIAL_DEGREE: Final = 4


class ExerciseTieRule(str, Enum):
    """Explicit decision when intrinsic and fitted continuation compare equal."""

    STRICTLY_GREATER = "strictly_greater"
    GREATER_OR_EQUAL = "greater_or_equal"


class InsufficientExerciseSupportPolicy(str, Enum):
    """Declared stopping rule when in-the-money rows cannot identify the polynomial basis."""

    FAIL = "fail"
    CONTINUE = "continue"


class Greek(str, Enum):
    """Pathwise sensitivities of the frozen-policy price."""

    DELTA = "delta"
    VEGA = "vega"
    RHO = "rho"
    DIVIDEND_YIELD = "dividend_yield"


GREEK_BITS: Final = {Greek.DELTA: 1, Greek.VEGA: 2, Greek.RHO: 4, Greek.DIVIDEND_YIELD: 8}
GREEK_ORDER: Final = (Greek.DELTA, Greek.VEGA, Greek.RHO, Greek.DIVIDEND_YIELD)


def greek_mask(greeks: tuple[Greek, ...]) -> int:
    if type(greeks) is not tuple or any(type(g) is not Greek for g in greeks):
        raise InputError("greeks must be an immutable tuple[Greek, ...]")
    mask = 0
    for g in greeks:
        mask |= GREEK_BITS[g]
    return mask


def _validate_random_streams(training: RandomStream, evaluation: RandomStream) -> None:
    if type(training) is not RandomStream:
        raise InputError("training_stream must be RandomStream")
    if type(evaluation) is not RandomStream:
        raise InputError("evaluation_stream must be RandomStream")
    if training.domain == evaluation.domain:
        msg = (
            "training_stream and evaluation_stream must use different domains; "
            f"both use {training.domain}"
        )
        raise InputError(msg)


@dataclass(frozen=True, slots=True)
class Lsm:
    """Complete simulation/regression identity for independent-policy LSM (reference defaults)."""

    training_paths: int
    evaluation_paths: int
    training_stream: RandomStream
    evaluation_stream: RandomStream
    insufficient_exercise_support: InsufficientExerciseSupportPolicy
    polynomial_degree: int = 4
    rank_tolerance: float = 1.0e-12
    condition_limit: float = 1.0e8
    tie_rule: ExerciseTieRule = ExerciseTieRule.STRICTLY_GREATER

    def __post_init__(self) -> None:
        for name, value in (("training_paths", self.training_paths),
                            ("evaluation_paths", self.evaluation_paths)):
            if type(value) is not int or not 1 <= value <= _MAX_PATHS:
                msg = f"{name} must be a plain int in [1, {_MAX_PATHS}]; observed {value!r}"
                raise InputError(msg)
        _validate_random_streams(self.training_stream, self.evaluation_stream)
        if type(self.insufficient_exercise_support) is not InsufficientExerciseSupportPolicy:
            msg = (
                "insufficient_exercise_support must be InsufficientExerciseSupportPolicy; "
                f"observed {type(self.insufficient_exercise_support).__name__}"
            )
            raise InputError(msg)
        if (type(self.polynomial_degree) is not int
                or not 0 <= self.polynomial_degree <= _MAX_POLYNOMIAL_DEGREE):
            msg = (
                "polynomial_degree must be a plain int in "
                f"[0, {_MAX_POLYNOMIAL_DEGREE}]; observed {self.polynomial_degree!r}"
            )
            raise InputError(msg)
        if (type(self.rank_tolerance) is not float or not isfinite(self.rank_tolerance)
                or not 0.0 < self.rank_tolerance < 1.0):
            msg = (
                "rank_tolerance must be a finite plain float strictly between 0 and 1; "
                f"observed {self.rank_tolerance!r}"
            )
            raise InputError(msg)
        maximum_condition = 1.0 / self.rank_tolerance
        if (type(self.condition_limit) is not float or not isfinite(self.condition_limit)
                or not 1.0 <= self.condition_limit <= maximum_condition):
            msg = (
                "condition_limit must be a finite plain float in "
                f"[1, {maximum_condition}]; observed {self.condition_limit!r}"
            )
            raise InputError(msg)
        if type(self.tie_rule) is not ExerciseTieRule:
            msg = f"tie_rule must be ExerciseTieRule; observed {type(self.tie_rule).__name__}"
            raise InputError(msg)


@dataclass(frozen=True, slots=True)
class OriginDecisionDiagnostics:
    """Common valuation-origin policy decision fitted from training continuation."""

    node: int
    exercise: bool
    estimated_continuation: float


@dataclass(frozen=True, slots=True)
class NoInTheMoneyPathsDiagnostics:
    """Training had no in-the-money state at this exercise node."""

    node: int


@dataclass(frozen=True, slots=True)
class InsufficientExerciseSupportDiagnostics:
    """Declared continuation at a node whose training population cannot identify the basis."""

    node: int
    in_the_money_paths: int
    required_paths: int


@dataclass(frozen=True, slots=True)
class RegressionFitDiagnostics:
    """Accepted continuation regression population and numerical diagnostics."""

    node: int
    in_the_money_paths: int
    observations: int
    columns: int
    rank: int
    condition_number: float
    predictor_mean: float
    predictor_scale: float


ExerciseFitDiagnostics = Union[
    OriginDecisionDiagnostics,
    NoInTheMoneyPathsDiagnostics,
    InsufficientExerciseSupportDiagnostics,
    RegressionFitDiagnostics
]


@dataclass(frozen=True, slots=True)
class LsmResult:
    """Independent mean payoff, sampling error, stopping counts, and fit diagnostics."""

    price: float
    paths: int
    sample_standard_deviation: float | None
    standard_error: float | None
    observation_nodes: tuple[int, ...]
    exercise_counts: tuple[int, ...]
    fit_diagnostics: tuple[ExerciseFitDiagnostics, ...]
    greeks: tuple[tuple[Greek, float, float], ...] = ()
    regression_coefficients: tuple[tuple[float, ...], ...] = ()
    # rule applications within the executor's near-tie window, backward pass and evaluation (no
    # counterpart in the reference's result)
    near_tie_training_decisions: int = 0
    near_tie_evaluation_decisions: int = 0

    def greek(self, name: Greek) -> float:
        for g, value, _se in self.greeks:
            if g is name:
                return value
        raise KeyError(name)

    def greek_standard_error(self, name: Greek) -> float:
        for g, _value, se in self.greeks:
            if g is name:
                return se
        raise KeyError(name)


def _diagnostic(d: FohipNodeDiagnostic) -> ExerciseFitDiagnostics:
    if d.kind == 0:
        return OriginDecisionDiagnostics(node=int(d.node), exercise=bool(d.exercise),
                                         estimated_continuation=float(d.estimated_continuation))
    if d.kind == 1:
        return NoInTheMoneyPathsDiagnostics(node=int(d.node))
    if d.kind == 2:
        return InsufficientExerciseSupportDiagnostics(
            node=int(d.node), in_the_money_paths=int(d.in_the_money_paths),
            required_paths=int(d.columns))
    if d.kind == 3:
        return RegressionFitDiagnostics(
            node=int(d.node), in_the_money_paths=int(d.in_the_money_paths),
            observations=int(d.observations), columns=int(d.columns), rank=int(d.rank),
            condition_number=float(d.condition_number), predictor_mean=float(d.predictor_mean),
            predictor_scale=float(d.predictor_scale))
    raise LsmError(f"native LSM result returned unsupported diagnostic kind {d.kind!r}")


def _validate_specs(*, initial_spot, strike, schedule, curve, dividend_yield_curve, volatility,
                    dividend_policy, executor, method) -> None:
    for name, value in (("initial_spot", initial_spot), ("strike", strike)):
        if type(value) is not float:
            raise InputError(f"{name} must be a plain float; observed {type(value).__name__}")
    checks = (("schedule", schedule, EventSchedule), ("curve", curve, LogDiscountCurve),
              ("dividend_yield_curve", dividend_yield_curve, CumulativeDividendYieldCurve),
              ("volatility", volatility, ConstantVolatility),
              ("dividend_policy", dividend_policy, DividendPolicy),
              ("executor", executor, HipExecutor), ("method", method, Lsm))
    for name, value, cls in checks:
        if type(value) is not cls:
            raise InputError(f"{name} must be {cls.__name__}; observed {type(value).__name__}")


def price_lsm_call(  # noqa: PLR0913 - financial inputs remain independently typed
    *,
    initial_spot: float,
    strike: float,
    schedule: EventSchedule,
    curve: LogDiscountCurve,
    dividend_yield_curve: CumulativeDividendYieldCurve,
    volatility: ConstantVolatility,
    dividend_policy: DividendPolicy,
    executor: HipExecutor,
    method: Lsm,
    greeks: tuple[Greek, ...] = (),
) -> LsmResult:
    """Fit a call stopping policy and evaluate it only on its independent path population.

    Exercise states are observed before any same-node cash dividend. Every estimator-defining
    numerical value is carried by ``method``; failures never reduce degree, add regularization,
    reuse training paths, or change the requested device or precision.

    Raises:
        InputError: If an explicit Python or financial specification is invalid.
        LsmError: If simulation, regression, support, or fixed-policy evaluation fails.
        ExecutionError: If the device resource fails.

    """
    _validate_specs(initial_spot=initial_spot, strike=strike, schedule=schedule, curve=curve,
                    dividend_yield_curve=dividend_yield_curve, volatility=volatility,
                    dividend_policy=dividend_policy, executor=executor, method=method)
    mask = greek_mask(greeks)
    f64 = lambda v: np.ascontiguousarray(v if len(v) else [0.0], dtype=np.float64)  # noqa: E731
    u32 = lambda v: np.ascontiguousarray(v if len(v) else [0], dtype=np.uint32)  # noqa: E731
    st, ct, cl = f64(schedule.times), f64(curve.times), f64(curve.log_discounts)
    yt, yq = f64(dividend_yield_curve.times), f64(dividend_yield_curve.cumulative_yields)
    dn, da = u32([d.node for d in schedule.dividends]), f64([d.amount for d in schedule.dividends])
    en = u32(list(schedule.exercise_nodes))
    req = FohipCallRequest(
        initial_spot, strike,
        ptr(st, C.c_double), len(schedule.times),
        ptr(ct, C.c_double), ptr(cl, C.c_double), len(curve.times),
        ptr(yt, C.c_double), ptr(yq, C.c_double), len(dividend_yield_curve.times),
        ptr(dn, C.c_uint32), ptr(da, C.c_double), len(schedule.dividends),
        ptr(en, C.c_uint32), len(schedule.exercise_nodes),
        schedule.expiry_node, volatility.annualized_decimal,
        0 if dividend_policy is DividendPolicy.LIQUIDATOR else 1,
        method.training_paths, method.evaluation_paths,
        method.training_stream.seed, method.training_stream.domain,
        method.evaluation_stream.seed, method.evaluation_stream.domain,
        method.polynomial_degree, method.rank_tolerance, method.condition_limit,
        0 if method.tie_rule is ExerciseTieRule.STRICTLY_GREATER else 1,
        0 if method.insufficient_exercise_support is InsufficientExerciseSupportPolicy.FAIL else 1,
        mask)
    obs = len(schedule.exercise_nodes) + 1
    nodes = np.zeros(obs, dtype=np.uint32)
    counts = np.zeros(obs, dtype=np.uint32)
    diags = (FohipNodeDiagnostic * max(obs - 1, 1))()
    res = FohipCallResult()
    res.observation_nodes = ptr(nodes, C.c_uint32)
    res.exercise_counts = ptr(counts, C.c_uint32)
    res.diagnostics = diags
    err = FohipError()
    code = lib().fohip_price_lsm_call(executor._handle.ptr, C.byref(req), C.byref(res),  # noqa: SLF001
                                      C.byref(err))
    raise_for(code, err)
    has_se = bool(res.has_standard_error)
    fit = tuple(_diagnostic(diags[r]) for r in range(obs - 1))
    coefficients = tuple(tuple(float(c) for c in diags[r].coefficients) for r in range(obs - 1))
    greek_values = tuple((g, float(res.greeks[i]), float(res.greek_standard_errors[i]))
                         for i, g in enumerate(GREEK_ORDER) if mask & GREEK_BITS[g])
    return LsmResult(
        price=float(res.price),
        paths=int(res.paths),
        sample_standard_deviation=float(res.sample_standard_deviation) if has_se else None,
        standard_error=float(res.standard_error) if has_se else None,
        observation_nodes=tuple(int(n) for n in nodes),
        exercise_counts=tuple(int(c) for c in counts),
        fit_diagnostics=fit,
        greeks=greek_values,
        regression_coefficients=coefficients,
        near_tie_training_decisions=int(res.near_tie_training_decisions),
        near_tie_evaluation_decisions=int(res.near_tie_evaluation_decisions),
    )


__all__ = [
    "ExerciseFitDiagnostics",
    "ExerciseTieRule",
    "Greek",
    "InsufficientExerciseSupportDiagnostics",
    "InsufficientExerciseSupportPolicy",
    "Lsm",
    "LsmError",
    "LsmResult",
    "NoInTheMoneyPathsDiagnostics",
    "OriginDecisionDiagnostics",
    "RegressionFitDiagnostics",
    "price_lsm_call",
]
