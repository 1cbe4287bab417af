"""ctypes view of libfohip.so (include/fohip.h) and the error-class mapping.

The library is loaded from this package directory only. A missing or unloadable library raises
``ExecutionError`` immediately: there is no CPU fallback anywhere in this package.
"""

from __future__ import annotations

import ctypes as C
import importlib.util
import os
from pathlib import Path

import numpy as np

LIB_PATH = Path(__file__).resolve().parent / "libfohip.so"


class InputError(ValueError):
    """Invalid explicit Python, financial or request specification (reference InputError)."""


class ExecutionError(RuntimeError):
    """Execution resource failure: no device, HIP error, allocation (reference ExecutionError)."""


class SimulationError(RuntimeError):
    """Path simulation failure (reference SimulationError)."""


class LsmError(RuntimeError):
    """Simulation, regression, support or evaluation failure of an LSM estimator."""


class FohipError(C.Structure):
    _fields_ = [
        ("code", C.c_int32),
        ("phase", C.c_int32),
        ("row", C.c_int64),
        ("node", C.c_int64),
        ("path", C.c_int64),
        ("index", C.c_int64),
        ("observed", C.c_double),
        ("limit", C.c_double),
        ("message", C.c_char * 512),
    ]


class FohipKernelStat(C.Structure):
    _fields_ = [("name", C.c_char * 48), ("launches", C.c_uint64), ("total_ms", C.c_double)]


class FohipTuning(C.Structure):
    _fields_ = [("eval_strike_tile", C.c_uint32), ("fit_cashflows", C.c_uint32),
                ("batch_schedule", C.c_uint32), ("batch_mib", C.c_uint32),
                ("fit_schedule", C.c_uint32)]


_dp = C.POINTER(C.c_double)
_u32p = C.POINTER(C.c_uint32)
_i64p = C.POINTER(C.c_int64)
_u8p = C.POINTER(C.c_uint8)
_u64p = C.POINTER(C.c_uint64)


class FohipCallRequest(C.Structure):
    _fields_ = [
        ("initial_spot", C.c_double),
        ("strike", C.c_double),
        ("schedule_times", _dp), ("schedule_nodes", C.c_size_t),
        ("curve_times", _dp), ("log_discounts", _dp), ("curve_nodes", C.c_size_t),
        ("yield_times", _dp), ("cumulative_yields", _dp), ("yield_nodes", C.c_size_t),
        ("dividend_nodes", _u32p), ("dividend_amounts", _dp), ("dividends", C.c_size_t),
        ("exercise_nodes", _u32p), ("exercise_count", C.c_size_t),
        ("expiry_node", C.c_uint32),
        ("volatility", C.c_double),
        ("dividend_policy", C.c_int32),
        ("training_paths", C.c_uint64), ("evaluation_paths", C.c_uint64),
        ("training_seed", C.c_uint64), ("training_domain", C.c_uint32),
        ("evaluation_seed", C.c_uint64), ("evaluation_domain", C.c_uint32),
        ("polynomial_degree", C.c_uint32),
        ("rank_tolerance", C.c_double), ("condition_limit", C.c_double),
        ("tie_rule", C.c_int32),
        ("insufficient_support", C.c_int32),
        ("greeks", C.c_uint32),
    ]


class FohipNodeDiagnostic(C.Structure):
    _fields_ = [
        ("node", C.c_uint32),
        ("kind", C.c_int32),
        ("exercise", C.c_int32),
        ("estimated_continuation", C.c_double),
        ("in_the_money_paths", C.c_uint32),
        ("observations", C.c_uint32),
        ("columns", C.c_uint32),
        ("rank", C.c_uint32),
        ("condition_number", C.c_double),
        ("predictor_mean", C.c_double),
        ("predictor_scale", C.c_double),
        ("alpha", C.c_double * 4),
        ("beta", C.c_double * 4),
        ("coefficients", C.c_double * 5),
    ]


class FohipCallResult(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("paths", C.c_uint32),
        ("has_standard_error", C.c_int32),
        ("sample_standard_deviation", C.c_double),
        ("standard_error", C.c_double),
        ("observation_rows", C.c_uint32),
        ("observation_nodes", _u32p),
        ("exercise_counts", _u32p),
        ("diagnostics", C.POINTER(FohipNodeDiagnostic)),
        ("greeks", C.c_double * 4),
        ("greek_standard_errors", C.c_double * 4),
        ("near_tie_training_decisions", C.c_uint64),
        ("near_tie_evaluation_decisions", C.c_uint64),
    ]


class FohipGroupedRequest(C.Structure):
    _fields_ = [
        ("group_offsets", _u32p), ("groups", C.c_size_t),
        ("initial_spots", _dp), ("maturities", _dp), ("rates", _dp), ("volatilities", _dp),
        ("strikes", _dp), ("rows", C.c_size_t),
        ("dividend_model", C.c_int32),
        ("dividend_yields", _dp),
        ("cash_offsets", _i64p), ("cash_times", _dp), ("cash_amounts", _dp),
        ("cash_events", C.c_size_t),
        ("cash_policy", C.c_int32), ("maturity_boundary", C.c_int32),
        ("mesh_kind", C.c_int32), ("base_intervals", C.c_uint32), ("origin_exercise", C.c_int32),
        ("exercise_offsets", _i64p), ("exercise_times", _dp), ("exercise_events", C.c_size_t),
        ("training_paths", C.c_uint64), ("evaluation_paths", C.c_uint64),
        ("training_path_counts", _u32p), ("evaluation_path_counts", _u32p),
        ("training_seed", C.c_uint64), ("evaluation_seed", C.c_uint64),
        ("training_domain_start", C.c_uint32), ("evaluation_domain_start", C.c_uint32),
        ("polynomial_degree", C.c_uint32),
        ("rank_tolerance", C.c_double), ("condition_limit", C.c_double),
        ("tie_rule", C.c_int32), ("insufficient_support", C.c_int32),
        ("greeks", C.c_uint32),
        ("escrowed_volatility", C.c_int32),
    ]


class FohipBatchResult(C.Structure):
    _fields_ = [
        ("prices", _dp), ("sample_standard_deviations", _dp), ("standard_errors", _dp),
        ("origin_decisions_fitted", _u8p), ("origin_exercised", _u8p),
        ("regression_nodes", _u32p), ("no_training_support_nodes", _u32p),
        ("insufficient_exercise_support_nodes", _u32p),
        ("minimum_in_the_money_training_paths", _u32p),
        ("maximum_regression_condition_numbers", _dp),
        ("early_exercise_paths", _u32p), ("expiry_paths", _u32p),
        ("delta", _dp), ("vega", _dp), ("rho", _dp), ("dividend_yield", _dp),
        ("delta_se", _dp), ("vega_se", _dp), ("rho_se", _dp), ("dividend_yield_se", _dp),
        ("near_tie_training_decisions", _u64p), ("near_tie_evaluation_decisions", _u64p),
    ]


class FohipPathPlan(C.Structure):
    _fields_ = [
        ("times", _dp), ("nodes", C.c_size_t),
        ("log_discounts", _dp),
        ("cumulative_yields", _dp),
        ("dividend_nodes", _u32p), ("dividend_amounts", _dp), ("dividends", C.c_size_t),
        ("exercise_nodes", _u32p), ("exercise_count", C.c_size_t),
        ("expiry_node", C.c_uint32),
        ("volatility", C.c_double),
        ("dividend_policy", C.c_int32),
    ]


# Every symbol declared in include/fohip.h (checked by the CPU test suite).
EXPORTED = (
    "fohip_abi_version", "fohip_device_count", "fohip_executor_create", "fohip_executor_destroy",
    "fohip_executor_info", "fohip_set_tuning", "fohip_set_near_tie_window", "fohip_price_lsm_call",
    "fohip_price_lsm_grouped",
    "fohip_grouped_diagnostics", "fohip_simulate_gbm_seeded", "fohip_simulate_gbm_seeded_device",
    "fohip_philox_blocks", "fohip_gbm_exp_probe",
    "fohip_stream_normals", "fohip_validate_time_grid", "fohip_validate_log_discount_curve",
    "fohip_validate_cumulative_yield_curve", "fohip_validate_event_schedule",
    "fohip_validate_constant_volatility", "fohip_validate_cash_dividend_amount",
    "fohip_validate_ragged_offsets", "fohip_validate_ragged_exercise_times",
    "fohip_validate_ragged_cash_dividends", "fohip_set_profiling", "fohip_get_profile",
    "fohip_reset_profile", "fohip_price_lsm_call_noalloc", "fohip_sweep_schedule",
)

_lib: C.CDLL | None = None


def _share_torch_hip_runtime() -> None:
    """Bind libfohip.so to the HIP runtime that torch ships, when torch is installed.

    torch bundles its own libamdhip64 under the soname of the system runtime (libamdhip64.so.7).
    Loaded first, libfohip.so would map the system copy, and torch would later map a second
    runtime that finds no device. Opening torch's copy first makes the dynamic loader resolve this
    library's dependency to it, so a process keeps one HIP runtime whichever package it imports
    first. torch itself is not imported.
    """
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.origin:
        return
    runtime = Path(spec.origin).parent / "lib" / "libamdhip64.so"
    if runtime.exists():
        try:
            C.CDLL(str(runtime), mode=os.RTLD_NOW | os.RTLD_LOCAL)
        except OSError:
            pass   # the system runtime then serves this library alone


def lib() -> C.CDLL:
    """Load libfohip.so from this package; raise ExecutionError when it is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        msg = (f"libfohip.so is not built at {LIB_PATH}; run finoptions-engine_amd/build_lib.py. "
               "There is no CPU fallback.")
        raise ExecutionError(msg)
    _share_torch_hip_runtime()
    try:
        handle = C.CDLL(str(LIB_PATH), mode=os.RTLD_NOW | os.RTLD_LOCAL)
    except OSError as exc:
        raise ExecutionError(f"libfohip.so failed to load: {exc}") from exc
    E = C.POINTER(FohipError)
    handle.fohip_executor_create.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_void_p), E]
    handle.fohip_executor_destroy.argtypes = [C.c_void_p]
    handle.fohip_executor_info.argtypes = [C.c_void_p, C.POINTER(C.c_int32), C.POINTER(C.c_int32),
                                           C.c_char_p, C.c_size_t]
    handle.fohip_set_tuning.argtypes = [C.c_void_p, C.POINTER(FohipTuning)]
    handle.fohip_set_near_tie_window.argtypes = [C.c_void_p, C.c_double]
    handle.fohip_sweep_schedule.argtypes = [_u32p, _u32p, C.c_uint32, C.c_uint32, _u32p, C.c_uint64,
                                            C.POINTER(C.c_uint64), _u32p]
    handle.fohip_price_lsm_call.argtypes = [C.c_void_p, C.POINTER(FohipCallRequest),
                                            C.POINTER(FohipCallResult), E]
    handle.fohip_price_lsm_call_noalloc.argtypes = [C.c_void_p, C.POINTER(FohipCallRequest),
                                                    C.POINTER(FohipCallResult), E]
    handle.fohip_price_lsm_grouped.argtypes = [C.c_void_p, C.POINTER(FohipGroupedRequest),
                                               C.POINTER(FohipBatchResult), E]
    handle.fohip_grouped_diagnostics.argtypes = [C.c_void_p, C.POINTER(FohipGroupedRequest),
                                                 C.c_uint32, C.POINTER(FohipNodeDiagnostic),
                                                 C.c_size_t, _u32p, E]
    handle.fohip_simulate_gbm_seeded.argtypes = [C.c_void_p, C.POINTER(FohipPathPlan), C.c_double,
                                                 C.c_uint64, C.c_uint64, C.c_uint32, _dp, E]
    handle.fohip_simulate_gbm_seeded_device.argtypes = [
        C.c_void_p, C.POINTER(FohipPathPlan), C.c_double, C.c_uint64, C.c_uint64, C.c_uint32,
        C.c_void_p, C.c_int32, C.c_void_p, E]
    handle.fohip_gbm_exp_probe.argtypes = [C.c_void_p, _dp, C.c_size_t, _dp, E]
    handle.fohip_philox_blocks.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                           C.c_uint32, C.c_uint32, _u32p, E]
    handle.fohip_stream_normals.argtypes = [C.c_void_p, C.c_uint64, C.c_uint32, C.c_uint32,
                                            C.c_uint32, C.c_uint32, _dp, _u32p, E]
    handle.fohip_validate_time_grid.argtypes = [_dp, C.c_size_t, E]
    handle.fohip_validate_log_discount_curve.argtypes = [_dp, _dp, C.c_size_t, E]
    handle.fohip_validate_cumulative_yield_curve.argtypes = [_dp, _dp, C.c_size_t, E]
    handle.fohip_validate_event_schedule.argtypes = [_dp, C.c_size_t, _u32p, _dp, C.c_size_t,
                                                     _u32p, C.c_size_t, C.c_uint32, E]
    handle.fohip_validate_constant_volatility.argtypes = [C.c_double, E]
    handle.fohip_validate_cash_dividend_amount.argtypes = [C.c_double, E]
    handle.fohip_validate_ragged_offsets.argtypes = [_i64p, C.c_size_t, C.c_size_t, E]
    handle.fohip_validate_ragged_exercise_times.argtypes = [_i64p, C.c_size_t, _dp, E]
    handle.fohip_validate_ragged_cash_dividends.argtypes = [_i64p, C.c_size_t, _dp, _dp, E]
    handle.fohip_device_count.argtypes = [C.POINTER(C.c_int32)]
    handle.fohip_set_profiling.argtypes = [C.c_void_p, C.c_int32]
    handle.fohip_get_profile.argtypes = [C.c_void_p, C.POINTER(FohipKernelStat), C.c_size_t,
                                         C.POINTER(C.c_size_t)]
    handle.fohip_reset_profile.argtypes = [C.c_void_p]
    _lib = handle
    return handle


def ptr(array: np.ndarray | None, ctype):
    if array is None:
        return C.cast(None, C.POINTER(ctype))
    return array.ctypes.data_as(C.POINTER(ctype))


INPUT_CODES = set(range(100, 200)) | set(range(500, 600))


def raise_for(code: int, err: FohipError, *, lsm: bool = True) -> None:
    """Map a failure code to the reference exception classes (lib.rs:57-64, :354-397)."""
    if code == 0:
        return
    message = err.message.decode(errors="replace") or f"fohip error {code}"
    if code in INPUT_CODES:
        exc: Exception = InputError(message)
    elif code >= 900:
        exc = ExecutionError(message)
    elif not lsm and 200 <= code < 300:
        exc = SimulationError(message)
    else:
        exc = LsmError(message)
    exc.code = code  # type: ignore[attr-defined]
    exc.phase = int(err.phase)  # type: ignore[attr-defined]
    exc.row = int(err.row)  # type: ignore[attr-defined]
    exc.node = int(err.node)  # type: ignore[attr-defined]
    exc.path = int(err.path)  # type: ignore[attr-defined]
    exc.index = int(err.index)  # type: ignore[attr-defined]
    exc.observed = float(err.observed)  # type: ignore[attr-defined]
    raise exc


def device_count() -> int:
    n = C.c_int32(0)
    lib().fohip_device_count(C.byref(n))
    return int(n.value)
