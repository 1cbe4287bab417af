"""MI355X-native Longstaff-Schwartz engine behind the finoptions pricing API.

Import surface mirrors python/finoptions/__init__.py for the LSM hot path: the same request
types, result types, error classes and ``price_lsm_call`` / ``price_lsm_call_grouped_batch``
entry points, with an explicit :class:`HipExecutor` in place of ``CpuExecutor``.
"""

from __future__ import annotations

from ._lib import ExecutionError, InputError, LsmError, SimulationError, device_count
from .batch_lsm import (
    BatchDividends,
    CashDividendBatch,
    CashDividendExerciseMesh,
    ContinuousDividendYieldBatch,
    EscrowedCashDividendBatch,
    EscrowedVolatilityConvention,
    ExerciseMesh,
    ExplicitExerciseMesh,
    LsmBatchResult,
    LsmStrikeGroups,
    OriginExercise,
    RaggedExerciseTimes,
    RaggedPathCounts,
    UniformExerciseMesh,
    grouped_fit_coefficients,
    price_lsm_call_batch,
    price_lsm_call_grouped_batch,
)
from .dividends import MaturityBoundary, RaggedCashDividends
from .lsm import (
    ExerciseFitDiagnostics,
    ExerciseTieRule,
    Greek,
    InsufficientExerciseSupportDiagnostics,
    InsufficientExerciseSupportPolicy,
    Lsm,
    LsmResult,
    NoInTheMoneyPathsDiagnostics,
    OriginDecisionDiagnostics,
    RegressionFitDiagnostics,
    price_lsm_call,
)
from .model import (
    CashDividend,
    ConstantVolatility,
    CumulativeDividendYieldCurve,
    DividendPolicy,
    EventSchedule,
    LogDiscountCurve,
)
from .simulation import (
    CpuExecutor,
    DeviceGbmPaths,
    GbmPaths,
    HipExecutor,
    HipPrecision,
    RandomStream,
    simulate_gbm_seeded,
    simulate_gbm_seeded_tensor,
)

__all__ = [
    "BatchDividends",
    "CashDividend",
    "CashDividendBatch",
    "CashDividendExerciseMesh",
    "ConstantVolatility",
    "ContinuousDividendYieldBatch",
    "EscrowedCashDividendBatch",
    "EscrowedVolatilityConvention",
    "CpuExecutor",
    "CumulativeDividendYieldCurve",
    "DividendPolicy",
    "EventSchedule",
    "ExecutionError",
    "ExerciseFitDiagnostics",
    "ExerciseMesh",
    "ExerciseTieRule",
    "ExplicitExerciseMesh",
    "DeviceGbmPaths",
    "GbmPaths",
    "Greek",
    "HipExecutor",
    "HipPrecision",
    "InputError",
    "InsufficientExerciseSupportDiagnostics",
    "InsufficientExerciseSupportPolicy",
    "LogDiscountCurve",
    "Lsm",
    "LsmBatchResult",
    "LsmError",
    "LsmResult",
    "LsmStrikeGroups",
    "MaturityBoundary",
    "NoInTheMoneyPathsDiagnostics",
    "OriginDecisionDiagnostics",
    "OriginExercise",
    "RaggedCashDividends",
    "RaggedExerciseTimes",
    "RaggedPathCounts",
    "RandomStream",
    "RegressionFitDiagnostics",
    "SimulationError",
    "UniformExerciseMesh",
    "device_count",
    "grouped_fit_coefficients",
    "price_lsm_call",
    "price_lsm_call_batch",
    "price_lsm_call_grouped_batch",
    "simulate_gbm_seeded",
    "simulate_gbm_seeded_tensor",
]
