"""Group-addressed and row-addressed batched LSM calls on HIP devices.

Mirrors python/finoptions/batch_lsm.py of the reference: ``LsmStrikeGroups``, the mesh and
dividend request types, ``LsmBatchResult`` and ``price_lsm_call_grouped_batch`` keep their names,
fields and checks; the executor is an explicit ``HipExecutor``. ``greeks`` adds pathwise
sensitivity columns under each strike's frozen stopping policy.

A row-addressed batch (``price_lsm_call_batch``) is the grouped batch with one strike per group:
row r uses training/evaluation domains start + r, exactly as the reference's row batch
(crates/finoptions-lsm/src/batch/input.rs:217-223).
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass
from enum import Enum
from typing import Final, Union

import numpy as np
import numpy.typing as npt

from ._lib import (
    FohipBatchResult,
    FohipError,
    FohipGroupedRequest,
    FohipNodeDiagnostic,
    InputError,
    LsmError,
    lib,
    ptr,
    raise_for,
)
from .dividends import MaturityBoundary, RaggedCashDividends, validate_vector
from .lsm import GREEK_BITS, ExerciseTieRule, Greek, InsufficientExerciseSupportPolicy, Lsm, greek_mask
from .model import DividendPolicy
from .simulation import HipExecutor

Float64Vector = npt.NDArray[np.float64]
Int64Vector = npt.NDArray[np.int64]
UInt32Vector = npt.NDArray[np.uint32]
UInt64Vector = npt.NDArray[np.uint64]
BoolVector = npt.NDArray[np.bool_]

_MAX_U32: Final = (2**32) - 1
_MIN_EVALUATION_PATHS: Final = 2


class EscrowedVolatilityConvention(str, Enum):
    """Which volatility the escrowed model's volatility column holds: the prepaid forward's own
    (PREPAID_FORWARD) or a spot volatility converted at the origin by S0 / F0
    (SPOT_MATCHED_AT_ORIGIN)."""

    PREPAID_FORWARD = "prepaid_forward"
    SPOT_MATCHED_AT_ORIGIN = "spot_matched_at_origin"


class OriginExercise(str, Enum):
    """Whether t = 0 is an exercise opportunity, decided once for all paths from the training
    cashflows (INCLUDED) or not (EXCLUDED)."""

    EXCLUDED = "excluded"
    INCLUDED = "included"


@dataclass(frozen=True, slots=True)
class UniformExerciseMesh:
    """Exercise grid T i / n for i = 0..n (n = ``base_intervals``), with every cash-dividend date
    inserted as its own node; every interior node is an exercise node."""

    base_intervals: int
    origin_exercise: OriginExercise

    def __post_init__(self) -> None:
        if type(self.base_intervals) is not int or not 1 <= self.base_intervals <= _MAX_U32:
            raise InputError(
                f"base_intervals must be a plain int in [1, {_MAX_U32}]; observed {self.base_intervals!r}")
        if type(self.origin_exercise) is not OriginExercise:
            raise InputError("origin_exercise must be OriginExercise")


@dataclass(frozen=True, slots=True)
class CashDividendExerciseMesh:
    """Exercise grid made of the cash-dividend dates alone (the state just before each payment),
    for cash dividends strictly before maturity and nonnegative rates."""


@dataclass(frozen=True, slots=True)
class RaggedExerciseTimes:
    """Per-row exercise times in Arrow list layout: int64 ``offsets`` (rows + 1 entries, possibly
    starting above 0) into one float64 ``times`` column; each row positive and strictly
    increasing. The arrays are borrowed, not copied."""

    offsets: Int64Vector
    times: Float64Vector

    def __post_init__(self) -> None:
        validate_vector(self.offsets, field="exercise offsets", dtype="int64")
        validate_vector(self.times, field="exercise times", dtype="float64")
        err = FohipError()
        code = lib().fohip_validate_ragged_offsets(ptr(self.offsets, C.c_int64), self.offsets.size,
                                                   self.times.size, C.byref(err))
        raise_for(code, err)
        code = lib().fohip_validate_ragged_exercise_times(
            ptr(self.offsets, C.c_int64), self.offsets.size, ptr(self.times, C.c_double), C.byref(err))
        raise_for(code, err)

    @property
    def rows(self) -> int:
        return int(self.offsets.size - 1)

    @property
    def events(self) -> int:
        return int(self.times.size)


@dataclass(frozen=True, slots=True)
class ExplicitExerciseMesh:
    """Exercise grid given explicitly per row (each time strictly before that row's maturity),
    united with its cash-dividend dates."""

    schedules: RaggedExerciseTimes
    origin_exercise: OriginExercise

    def __post_init__(self) -> None:
        if type(self.schedules) is not RaggedExerciseTimes:
            raise InputError("schedules must be RaggedExerciseTimes")
        if type(self.origin_exercise) is not OriginExercise:
            raise InputError("origin_exercise must be OriginExercise")


@dataclass(frozen=True, slots=True)
class ContinuousDividendYieldBatch:
    """Continuous proportional dividend yield q >= 0 per group: Q(t) = q t on its grid."""

    annualized_decimals: Float64Vector

    def __post_init__(self) -> None:
        validate_vector(self.annualized_decimals, field="annualized_decimals", dtype="float64")


@dataclass(frozen=True, slots=True)
class CashDividendBatch:
    """Cash dividends per group (Arrow layout), the policy for a dividend larger than the spot,
    and whether a payment exactly at maturity is admitted."""

    schedules: RaggedCashDividends
    policy: DividendPolicy
    maturity_boundary: MaturityBoundary

    def __post_init__(self) -> None:
        if type(self.schedules) is not RaggedCashDividends:
            raise InputError("schedules must be RaggedCashDividends")
        if type(self.policy) is not DividendPolicy:
            raise InputError("policy must be DividendPolicy")
        if type(self.maturity_boundary) is not MaturityBoundary:
            raise InputError("maturity_boundary must be MaturityBoundary")


@dataclass(frozen=True, slots=True)
class EscrowedCashDividendBatch:
    """Escrowed cash dividends: the stock is a GBM prepaid forward F plus the discounted value of
    the dividends still to be paid.

    An escrowed-dividend approximation, distinct from the exact spot cash-jump model of
    :class:`CashDividendBatch`: the simulated state is the prepaid forward F (no dividends), started
    at F0 = S0 - sum_k D_k exp(-r t_k), and an exercise date observes F + the prepaid value of the
    cash still to come (that date's amount included). ``volatility_convention`` states whether the
    volatility column belongs to F already or is a spot volatility to be rescaled by S0 / F0
    (python/finoptions/batch_lsm.py:171-195 of the reference).
    """

    schedules: RaggedCashDividends
    volatility_convention: EscrowedVolatilityConvention
    maturity_boundary: MaturityBoundary

    def __post_init__(self) -> None:
        if type(self.schedules) is not RaggedCashDividends:
            raise InputError("schedules must be RaggedCashDividends")
        if type(self.volatility_convention) is not EscrowedVolatilityConvention:
            raise InputError("volatility_convention must be EscrowedVolatilityConvention")
        if type(self.maturity_boundary) is not MaturityBoundary:
            raise InputError("maturity_boundary must be MaturityBoundary")


@dataclass(frozen=True, slots=True)
class RaggedPathCounts:
    """Per-group (per-contract for a row batch) training and evaluation path counts.

    An extension for ragged batches: the reference's ``Lsm`` carries one path count for every row.
    When given, these counts replace ``method.training_paths`` / ``method.evaluation_paths``.
    """

    training: UInt32Vector
    evaluation: UInt32Vector

    def __post_init__(self) -> None:
        validate_vector(self.training, field="training path counts", dtype="uint32")
        validate_vector(self.evaluation, field="evaluation path counts", dtype="uint32")
        if self.training.size != self.evaluation.size:
            raise InputError("training and evaluation path counts must have one entry per group")
        if np.any(self.training == 0):
            raise InputError("training path counts must be positive")
        if np.any(self.evaluation < _MIN_EVALUATION_PATHS):
            raise InputError("ragged batches require at least two evaluation paths per group")


BatchDividends = Union[ContinuousDividendYieldBatch, CashDividendBatch, EscrowedCashDividendBatch]
_DIVIDEND_TYPES: Final = (ContinuousDividendYieldBatch, CashDividendBatch, EscrowedCashDividendBatch)
_DIVIDEND_TYPE_MESSAGE: Final = (
    "dividends must be ContinuousDividendYieldBatch, CashDividendBatch, or EscrowedCashDividendBatch")
ExerciseMesh = Union[UniformExerciseMesh, CashDividendExerciseMesh, ExplicitExerciseMesh]


@dataclass(frozen=True, slots=True)
class LsmStrikeGroups:
    """Market states (spot, maturity, rate, volatility, dividends) with their strike rows:
    uint32 ``group_offsets`` (groups + 1, from 0 to the strike count) partition ``strikes`` into
    the rows of each state, which share one simulated path population."""

    group_offsets: UInt32Vector
    initial_spots: Float64Vector
    strikes: Float64Vector
    maturities: Float64Vector
    rates: Float64Vector
    volatilities: Float64Vector
    dividends: BatchDividends

    def __post_init__(self) -> None:
        validate_vector(self.group_offsets, field="group_offsets", dtype="uint32")
        for name, values in (("initial_spots", self.initial_spots), ("strikes", self.strikes),
                             ("maturities", self.maturities), ("rates", self.rates),
                             ("volatilities", self.volatilities)):
            validate_vector(values, field=name, dtype="float64")
        if type(self.dividends) not in _DIVIDEND_TYPES:
            raise InputError(_DIVIDEND_TYPE_MESSAGE)


_COLUMNS: Final = (
    ("prices", "float64"), ("sample_standard_deviations", "float64"),
    ("standard_errors", "float64"), ("origin_decisions_fitted", "bool"),
    ("origin_exercised", "bool"), ("regression_nodes", "uint32"),
    ("no_training_support_nodes", "uint32"), ("insufficient_exercise_support_nodes", "uint32"),
    ("minimum_in_the_money_training_paths", "uint32"),
    ("maximum_regression_condition_numbers", "float64"), ("early_exercise_paths", "uint32"),
    ("expiry_paths", "uint32"),
)


@dataclass(frozen=True, slots=True, eq=False)
class LsmBatchResult:
    """Immutable row-ordered prices, sampling error, policy support, stopping counts, Greeks.

    ``near_tie_training_decisions`` and ``near_tie_evaluation_decisions`` count, per row, the rule
    applications of the backward pass and of the evaluation that are ties: by default exact ties,
    intrinsic == continuation, which the tie rule settles; with a window
    (``HipExecutor.set_near_tie_window``) every decision within it. They have no counterpart in
    the reference's result and are filled by every HIP pricing.
    """

    prices: Float64Vector
    sample_standard_deviations: Float64Vector
    standard_errors: Float64Vector
    origin_decisions_fitted: BoolVector
    origin_exercised: BoolVector
    regression_nodes: UInt32Vector
    no_training_support_nodes: UInt32Vector
    insufficient_exercise_support_nodes: UInt32Vector
    minimum_in_the_money_training_paths: UInt32Vector
    maximum_regression_condition_numbers: Float64Vector
    early_exercise_paths: UInt32Vector
    expiry_paths: UInt32Vector
    greeks: dict[Greek, Float64Vector] | None = None
    greek_standard_errors: dict[Greek, Float64Vector] | None = None
    near_tie_training_decisions: UInt64Vector | None = None
    near_tie_evaluation_decisions: UInt64Vector | None = None

    def __post_init__(self) -> None:
        rows = int(self.prices.size)
        for name, dtype in _COLUMNS:
            values = getattr(self, name)
            validate_vector(values, field=name, dtype=dtype)
            if values.size != rows:
                raise InputError(f"{name} has {values.size} rows; expected {rows}")
            if values.flags.writeable:
                raise InputError(f"{name} must be immutable")

    @property
    def rows(self) -> int:
        return int(self.prices.size)


def _mesh_fields(mesh: ExerciseMesh, groups: int) -> dict:
    if type(mesh) is UniformExerciseMesh:
        return {"mesh_kind": 0, "base_intervals": mesh.base_intervals,
                "origin_exercise": 0 if mesh.origin_exercise is OriginExercise.EXCLUDED else 1}
    if type(mesh) is CashDividendExerciseMesh:
        return {"mesh_kind": 1, "base_intervals": 1, "origin_exercise": 0}
    if type(mesh) is ExplicitExerciseMesh:
        if mesh.schedules.rows != groups:
            raise InputError(
                f"explicit exercise schedules have {mesh.schedules.rows} rows; expected {groups}")
        return {"mesh_kind": 2, "base_intervals": 0,
                "origin_exercise": 0 if mesh.origin_exercise is OriginExercise.EXCLUDED else 1,
                "exercise": mesh.schedules}
    raise InputError(
        "mesh must be UniformExerciseMesh, CashDividendExerciseMesh, or ExplicitExerciseMesh")


def _dividend_fields(req: FohipGroupedRequest, dividends: BatchDividends, groups: int) -> None:
    """Dividend model fields of the C request (the three reference variants,
    python/finoptions/batch_lsm.py:596-627)."""
    if type(dividends) is ContinuousDividendYieldBatch:
        if dividends.annualized_decimals.size != groups:
            raise InputError(
                f"dividend_model has {dividends.annualized_decimals.size} rows; expected {groups}")
        req.dividend_model = 0
        req.dividend_yields = ptr(dividends.annualized_decimals, C.c_double)
        return
    if type(dividends) not in (CashDividendBatch, EscrowedCashDividendBatch):
        raise InputError(_DIVIDEND_TYPE_MESSAGE)
    sched = dividends.schedules
    if sched.rows != groups:
        raise InputError(f"dividend_model has {sched.rows} rows; expected {groups}")
    req.cash_offsets = ptr(sched.offsets, C.c_int64)
    req.cash_times = ptr(sched.times, C.c_double)
    req.cash_amounts = ptr(sched.amounts, C.c_double)
    req.cash_events = sched.times.size
    req.maturity_boundary = (0 if dividends.maturity_boundary is MaturityBoundary.STRICTLY_BEFORE
                             else 1)
    if type(dividends) is CashDividendBatch:
        req.dividend_model = 1
        req.cash_policy = 0 if dividends.policy is DividendPolicy.LIQUIDATOR else 1
    else:
        req.dividend_model = 2
        req.escrowed_volatility = (
            0 if dividends.volatility_convention is EscrowedVolatilityConvention.PREPAID_FORWARD
            else 1)


def _price(group_offsets: np.ndarray, initial_spots, strikes, maturities, rates, volatilities,
           dividends: BatchDividends, mesh: ExerciseMesh, executor: HipExecutor, method: Lsm,
           greeks: tuple[Greek, ...], *, what: str,
           path_counts: RaggedPathCounts | None = None, tune: bool = True) -> LsmBatchResult:
    if type(executor) is not HipExecutor:
        raise InputError("executor must be HipExecutor")
    if type(method) is not Lsm:
        raise InputError("method must be Lsm")
    if method.evaluation_paths < _MIN_EVALUATION_PATHS:
        raise InputError(f"{what} requires at least two independent evaluation paths")
    mask = greek_mask(greeks)
    groups = int(initial_spots.size)
    for name, values in (("maturities", maturities), ("rates", rates),
                         ("volatilities", volatilities)):
        if values.size != groups:
            raise InputError(f"{name} has {values.size} rows; expected {groups}")
    mf = _mesh_fields(mesh, groups)
    if mf["mesh_kind"] == 1 and type(dividends) is ContinuousDividendYieldBatch:
        raise InputError("CashDividendExerciseMesh requires cash-dividend inputs")
    if tune and executor.autotune and groups > 0:
        from .autotune import tune_request

        def sample_price(**cols):
            return _price(cols["group_offsets"], cols["initial_spots"], cols["strikes"],
                          cols["maturities"], cols["rates"], cols["volatilities"], cols["dividends"],
                          cols["mesh"], executor, method, greeks, what=what,
                          path_counts=cols["path_counts"], tune=False)

        tune_request(executor, sample_price, group_offsets=group_offsets, initial_spots=initial_spots,
                     strikes=strikes, maturities=maturities, rates=rates, volatilities=volatilities,
                     dividends=dividends, mesh=mesh, method=method, greeks=greeks,
                     path_counts=path_counts)
    rows = int(strikes.size)
    keep: list = []
    req = FohipGroupedRequest()
    req.group_offsets = ptr(group_offsets, C.c_uint32)
    req.groups = groups
    req.initial_spots = ptr(initial_spots, C.c_double)
    req.maturities = ptr(maturities, C.c_double)
    req.rates = ptr(rates, C.c_double)
    req.volatilities = ptr(volatilities, C.c_double)
    req.strikes = ptr(strikes, C.c_double)
    req.rows = rows
    _dividend_fields(req, dividends, groups)
    req.mesh_kind = mf["mesh_kind"]
    req.base_intervals = mf["base_intervals"]
    req.origin_exercise = mf["origin_exercise"]
    if "exercise" in mf:
        req.exercise_offsets = ptr(mf["exercise"].offsets, C.c_int64)
        req.exercise_times = ptr(mf["exercise"].times, C.c_double)
        req.exercise_events = mf["exercise"].times.size
    req.training_paths = method.training_paths
    req.evaluation_paths = method.evaluation_paths
    if path_counts is not None:
        if type(path_counts) is not RaggedPathCounts:
            raise InputError("path_counts must be RaggedPathCounts")
        if path_counts.training.size != groups:
            raise InputError(f"path_counts has {path_counts.training.size} groups; expected {groups}")
        req.training_path_counts = ptr(path_counts.training, C.c_uint32)
        req.evaluation_path_counts = ptr(path_counts.evaluation, C.c_uint32)
    req.training_seed = method.training_stream.seed
    req.evaluation_seed = method.evaluation_stream.seed
    req.training_domain_start = method.training_stream.domain
    req.evaluation_domain_start = method.evaluation_stream.domain
    req.polynomial_degree = method.polynomial_degree
    req.rank_tolerance = method.rank_tolerance
    req.condition_limit = method.condition_limit
    req.tie_rule = 0 if method.tie_rule is ExerciseTieRule.STRICTLY_GREATER else 1
    req.insufficient_support = (
        0 if method.insufficient_exercise_support is InsufficientExerciseSupportPolicy.FAIL else 1)
    req.greeks = mask
    cols = {}
    for name, dtype in _COLUMNS:
        cols[name] = np.zeros(rows, dtype=np.uint8 if dtype == "bool" else dtype)
    res = FohipBatchResult()
    ctype = {"float64": C.c_double, "uint32": C.c_uint32, "bool": C.c_uint8}
    for name, dtype in _COLUMNS:
        setattr(res, name, ptr(cols[name], ctype[dtype]))
    greek_cols: dict[Greek, np.ndarray] = {}
    greek_se: dict[Greek, np.ndarray] = {}
    for g in (Greek.DELTA, Greek.VEGA, Greek.RHO, Greek.DIVIDEND_YIELD):
        if mask & GREEK_BITS[g]:
            greek_cols[g] = np.zeros(rows)
            greek_se[g] = np.zeros(rows)
            setattr(res, g.value, ptr(greek_cols[g], C.c_double))
            setattr(res, g.value + "_se", ptr(greek_se[g], C.c_double))
    ties = {name: np.zeros(rows, dtype=np.uint64)
            for name in ("near_tie_training_decisions", "near_tie_evaluation_decisions")}
    for name, values in ties.items():
        setattr(res, name, ptr(values, C.c_uint64))
    keep.append((cols, greek_cols, greek_se, ties))
    err = FohipError()
    code = lib().fohip_price_lsm_grouped(executor._handle.ptr, C.byref(req), C.byref(res),  # noqa: SLF001
                                         C.byref(err))
    raise_for(code, err)
    for name, dtype in _COLUMNS:
        if dtype == "bool":
            cols[name] = cols[name].astype(np.bool_)
        cols[name].flags.writeable = False
    for d in (greek_cols, greek_se, ties):
        for v in d.values():
            v.flags.writeable = False
    result = LsmBatchResult(**cols, greeks=greek_cols or None, greek_standard_errors=greek_se or None, **ties)
    stopping = result.early_exercise_paths.astype(np.uint64) + result.expiry_paths.astype(np.uint64)
    if path_counts is None:
        expected = np.full(rows, method.evaluation_paths, dtype=np.uint64)
    else:
        expected = np.repeat(path_counts.evaluation.astype(np.uint64),
                             np.diff(group_offsets.astype(np.int64)))
    if np.any(stopping != expected):
        raise LsmError(f"native {what} stopping counts do not partition evaluation paths")
    return result


def price_lsm_call_grouped_batch(
    *,
    groups: LsmStrikeGroups,
    mesh: ExerciseMesh,
    executor: HipExecutor,
    method: Lsm,
    greeks: tuple[Greek, ...] = (),
    path_counts: RaggedPathCounts | None = None,
) -> LsmBatchResult:
    """Price every strike row of every group; each group's state is simulated once and shared.

    A strike row has its own stopping policy, fitted on its group's training paths and valued on
    the group's evaluation paths, so the strikes of a group see common random numbers. Group g
    draws from domains ``training_stream.domain + g`` and ``evaluation_stream.domain + g``.

    Raises:
        InputError: If a runtime type, array layout, compact boundary, financial value, mesh, or
            group-addressed random interval is invalid.
        LsmError: If simulation, regression, policy support, or evaluation fails.
        ExecutionError: If the device resource fails.

    """
    if type(groups) is not LsmStrikeGroups:
        raise InputError("groups must be LsmStrikeGroups")
    if type(mesh) not in (UniformExerciseMesh, CashDividendExerciseMesh, ExplicitExerciseMesh):
        raise InputError(
            "mesh must be UniformExerciseMesh, CashDividendExerciseMesh, or ExplicitExerciseMesh")
    return _price(groups.group_offsets, groups.initial_spots, groups.strikes, groups.maturities,
                  groups.rates, groups.volatilities, groups.dividends, mesh, executor, method,
                  greeks, what="grouped batched LSM", path_counts=path_counts)


def price_lsm_call_batch(  # noqa: PLR0913 - each financial column is independently typed
    *,
    initial_spots: Float64Vector,
    strikes: Float64Vector,
    maturities: Float64Vector,
    rates: Float64Vector,
    volatilities: Float64Vector,
    dividends: BatchDividends,
    mesh: ExerciseMesh,
    executor: HipExecutor,
    method: Lsm,
    greeks: tuple[Greek, ...] = (),
    path_counts: RaggedPathCounts | None = None,
) -> LsmBatchResult:
    """Price constant-parameter Bermudan calls, one simulated state per contract row.

    With ``path_counts`` each contract has its own training/evaluation path count (a ragged
    batch); all contracts are still priced in one request over shared device launches.
    """
    for name, values in (("initial_spots", initial_spots), ("strikes", strikes),
                         ("maturities", maturities), ("rates", rates),
                         ("volatilities", volatilities)):
        validate_vector(values, field=name, dtype="float64")
    if strikes.size != initial_spots.size:
        raise InputError(f"strikes has {strikes.size} rows; expected {initial_spots.size}")
    offsets = np.arange(initial_spots.size + 1, dtype=np.uint32)
    return _price(offsets, initial_spots, strikes, maturities, rates, volatilities, dividends, mesh,
                  executor, method, greeks, what="batched LSM", path_counts=path_counts)


def grouped_fit_coefficients(*, groups: LsmStrikeGroups, mesh: ExerciseMesh,
                             executor: HipExecutor, method: Lsm, group: int) -> np.ndarray:
    """Fitted rule of every (strike, exercise row) of one group, shape [strikes, rows, 17].

    Columns: mean, scale, alpha[4], beta[4], coef[5], condition, kind (0 origin, 1 no support,
    2 insufficient support, 3 regression); a diagnostic view used by parity tests.
    """
    lo, hi = int(groups.group_offsets[group]), int(groups.group_offsets[group + 1])
    req, _keep = _request_for(groups, mesh, method)
    n_rows = hi - lo
    rows = C.c_uint32(0)
    cap = max(n_rows, 1) * 4096
    buf = (FohipNodeDiagnostic * cap)()
    err = FohipError()
    code = lib().fohip_grouped_diagnostics(executor._handle.ptr, C.byref(req), group, buf, cap,  # noqa: SLF001
                                           C.byref(rows), C.byref(err))
    raise_for(code, err)
    per = int(rows.value)
    out = np.zeros((n_rows, per, 17))
    for s in range(n_rows):
        for r in range(per):
            d = buf[s * per + r]
            out[s, r] = [d.predictor_mean, d.predictor_scale, *d.alpha, *d.beta, *d.coefficients,
                         d.condition_number, float(d.kind)]
    return out


def _request_for(groups: LsmStrikeGroups, mesh: ExerciseMesh, method: Lsm):
    keep = []
    req = FohipGroupedRequest()
    req.group_offsets = ptr(groups.group_offsets, C.c_uint32)
    req.groups = groups.initial_spots.size
    req.initial_spots = ptr(groups.initial_spots, C.c_double)
    req.maturities = ptr(groups.maturities, C.c_double)
    req.rates = ptr(groups.rates, C.c_double)
    req.volatilities = ptr(groups.volatilities, C.c_double)
    req.strikes = ptr(groups.strikes, C.c_double)
    req.rows = groups.strikes.size
    _dividend_fields(req, groups.dividends, int(groups.initial_spots.size))
    mf = _mesh_fields(mesh, int(groups.initial_spots.size))
    req.mesh_kind = mf["mesh_kind"]
    req.base_intervals = mf["base_intervals"]
    req.origin_exercise = mf["origin_exercise"]
    if "exercise" in mf:
        req.exercise_offsets = ptr(mf["exercise"].offsets, C.c_int64)
        req.exercise_times = ptr(mf["exercise"].times, C.c_double)
        req.exercise_events = mf["exercise"].times.size
    req.training_paths = method.training_paths
    req.evaluation_paths = method.evaluation_paths
    req.training_seed = method.training_stream.seed
    req.evaluation_seed = method.evaluation_stream.seed
    req.training_domain_start = method.training_stream.domain
    req.evaluation_domain_start = method.evaluation_stream.domain
    req.polynomial_degree = method.polynomial_degree
    req.rank_tolerance = method.rank_tolerance
    req.condition_limit = method.condition_limit
    req.tie_rule = 0 if method.tie_rule is ExerciseTieRule.STRICTLY_GREATER else 1
    req.insufficient_support = (
        0 if method.insufficient_exercise_support is InsufficientExerciseSupportPolicy.FAIL else 1)
    return req, keep


__all__ = [
    "BatchDividends",
    "CashDividendBatch",
    "CashDividendExerciseMesh",
    "ContinuousDividendYieldBatch",
    "EscrowedCashDividendBatch",
    "EscrowedVolatilityConvention",
    "ExerciseMesh",
    "ExplicitExerciseMesh",
    "LsmBatchResult",
    "LsmStrikeGroups",
    "OriginExercise",
    "RaggedExerciseTimes",
    "RaggedPathCounts",
    "UniformExerciseMesh",
    "grouped_fit_coefficients",
    "price_lsm_call_batch",
    "price_lsm_call_grouped_batch",
]
