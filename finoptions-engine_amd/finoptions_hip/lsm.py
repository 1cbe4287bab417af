"""Longstaff-Schwartz pricing of one scheduled Bermudan call on a HIP device.

The public names, fields, defaults and error messages are those of the reference's
python/finoptions/lsm.py (``Lsm``, the per-node diagnostics, ``LsmResult``, ``price_lsm_call``), so
requests port unchanged; the executor is an explicit ``HipExecutor`` and ``greeks`` requests
pathwise sensitivities under the frozen policy. Request checks are the rule tables below.
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
from . import _rules as R
from .simulation import HipExecutor, RandomStream

_MAX_PATHS: Final = (2**32) - 1
_MAX_POLYNOMIAL_DEGREE: Final = 4


class ExerciseTieRule(str, Enum):
    """How an exercise decision treats intrinsic == continuation: STRICTLY_GREATER holds the
    option, GREATER_OR_EQUAL exercises it."""

    STRICTLY_GREATER = "strictly_greater"
    GREATER_OR_EQUAL = "greater_or_equal"


class InsufficientExerciseSupportPolicy(str, Enum):
    """What a node with fewer in-the-money training paths than basis columns does: FAIL raises,
    CONTINUE records the node and never exercises there."""

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


def _method_rows(m: "Lsm") -> list:
    """The checks of an Lsm method in reporting order: path counts, the two streams and their
    distinct domains, the support policy, degree, rank tolerance, condition limit (bounded by
    1 / rank tolerance) and tie rule."""
    path_count = R.int_in(1, _MAX_PATHS)
    streams = (m.training_stream, m.evaluation_stream)
    rows = [("training_paths", m.training_paths, path_count),
            ("evaluation_paths", m.evaluation_paths, path_count),
            ("training_stream", m.training_stream,
             (R.exact_type(RandomStream, "{field} must be RandomStream"),)),
            ("evaluation_stream", m.evaluation_stream,
             (R.exact_type(RandomStream, "{field} must be RandomStream"),)),
            ("streams", streams,
             (R.Lazy(lambda st: st[0].domain != st[1].domain,
                     lambda st: "training_stream and evaluation_stream must use different "
                                f"domains; both use {st[0].domain}"),)),
            ("insufficient_exercise_support", m.insufficient_exercise_support,
             R.instance_of(InsufficientExerciseSupportPolicy)),
            ("polynomial_degree", m.polynomial_degree, R.int_in(0, _MAX_POLYNOMIAL_DEGREE)),
            ("rank_tolerance", m.rank_tolerance,
             (R.Check(lambda v: type(v) is float and isfinite(v) and 0.0 < v < 1.0,
                      "{field} must be a finite plain float strictly between 0 and 1; "
                      "observed {value!r}"),))]
    # rows are checked in order, so the condition limit's bound 1 / rank_tolerance is formed only
    # once the tolerance has passed its own row
    rows += [("condition_limit", m.condition_limit,
              (R.Lazy(lambda v: type(v) is float and isfinite(v) and 1.0 <= v <= 1.0 / m.rank_tolerance,
                      lambda v: "condition_limit must be a finite plain float in "
                                f"[1, {1.0 / m.rank_tolerance}]; observed {v!r}"),)),
             ("tie_rule", m.tie_rule, R.instance_of(ExerciseTieRule))]
    return rows


@dataclass(frozen=True, slots=True)
class Lsm:
    """The estimator: path counts and streams of the training and evaluation populations, the
    support policy, the polynomial basis degree with its rank tolerance and condition limit, and
    the tie rule. The two streams must use different domains so that the evaluation paths are
    independent of the paths the policy was fitted on."""

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
        R.enforce(_method_rows(self))


@dataclass(frozen=True, slots=True)
class OriginDecisionDiagnostics:
    """The valuation origin's decision: exercise when the intrinsic value beats the mean
    discounted training cashflow ``estimated_continuation``."""

    node: int
    exercise: bool
    estimated_continuation: float


@dataclass(frozen=True, slots=True)
class NoInTheMoneyPathsDiagnostics:
    """An exercise node at which no training path was in the money (no rule was fitted)."""

    node: int


@dataclass(frozen=True, slots=True)
class InsufficientExerciseSupportDiagnostics:
    """An exercise node whose in-the-money training paths were fewer than the basis columns, under
    the CONTINUE policy: no exercise there."""

    node: int
    in_the_money_paths: int
    required_paths: int


@dataclass(frozen=True, slots=True)
class RegressionFitDiagnostics:
    """A fitted continuation regression: its in-the-money population, design shape and numerical
    rank, condition number, and the moneyness standardization (mean, scale)."""

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
    """One pricing: the mean discounted payoff over the evaluation paths with its sample standard
    deviation and standard error (None for a single path), the observation nodes with the
    number of paths stopping at each, one diagnostic per exercise node, and the requested
    Greeks as (greek, value, standard error)."""

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


def _call_rows(initial_spot, strike, schedule, curve, dividend_yield_curve, volatility,
               dividend_policy, executor, method) -> list:
    """Exact runtime types of a single-call request, in the reference's reporting order."""
    return [("initial_spot", initial_spot, R.PLAIN_FLOAT), ("strike", strike, R.PLAIN_FLOAT),
            ("schedule", schedule, R.instance_of(EventSchedule)),
            ("curve", curve, R.instance_of(LogDiscountCurve)),
            ("dividend_yield_curve", dividend_yield_curve,
             R.instance_of(CumulativeDividendYieldCurve)),
            ("volatility", volatility, R.instance_of(ConstantVolatility)),
            ("dividend_policy", dividend_policy, R.instance_of(DividendPolicy)),
            ("executor", executor, R.instance_of(HipExecutor)),
            ("method", method, R.instance_of(Lsm))]


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
    allocation_free: bool = False,
) -> LsmResult:
    """Price a Bermudan call on the HIP device: fit the stopping policy by backward induction on
    the training paths, then value it on the independent evaluation paths.

    The request is taken as given: a failed fit raises instead of lowering the degree or
    regularizing, the evaluation never reuses training paths, and neither the device nor the
    executor's method is changed. ``greeks`` adds pathwise sensitivities of the price under the
    fitted (frozen) policy. ``allocation_free`` prices through fohip_price_lsm_call_noalloc, whose
    device buffers come from a fixed per-device region instead of device allocations (bitwise the
    same results; requests too large for the region raise ExecutionError).

    Raises:
        InputError: If a runtime type or a financial input is invalid.
        LsmError: If simulation, a regression, the exercise support or the evaluation fails.
        ExecutionError: If the device fails.

    """
    R.enforce(_call_rows(initial_spot, strike, schedule, curve, dividend_yield_curve, volatility,
                         dividend_policy, executor, method))
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
    if type(allocation_free) is not bool:
        raise InputError("allocation_free must be a plain bool")
    entry = lib().fohip_price_lsm_call_noalloc if allocation_free else lib().fohip_price_lsm_call
    code = entry(executor._handle.ptr, C.byref(req), C.byref(res), C.byref(err))  # noqa: SLF001
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
