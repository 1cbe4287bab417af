"""Explicit HIP execution resource, random streams and seeded path export.

``HipExecutor`` is the MI355X counterpart of the reference ``CpuExecutor``
(python/finoptions/simulation.py:40-62): an explicit, immutable resource naming one device and one
numerical method. Construction fails when the device is absent; nothing falls back to the CPU.
"""

from __future__ import annotations

import ctypes as C
import weakref
from dataclasses import dataclass, field
from enum import Enum
from itertools import pairwise
from typing import Final

import numpy as np
import numpy.typing as npt

from ._lib import (
    ExecutionError,
    FohipError,
    FohipKernelStat,
    FohipPathPlan,
    FohipTuning,
    InputError,
    SimulationError,
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

Float64Matrix = npt.NDArray[np.float64]
_MAX_SEED: Final = (2**64) - 1
_MAX_DOMAIN: Final = (2**32) - 1
_MAX_PATHS: Final = (2**32) - 1


class HipPrecision(str, Enum):
    """Numerical method of a HIP executor.

    ``F64``: the reference's algorithm and arithmetic in binary64 (a Welford pass and d + 1
    Stieltjes passes per exercise node); matches the reference CPU path.
    ``FUSED_F64``: one fused pass per node (pivoted power sums, moment-form Stieltjes solve) over
    binary64 spots and cashflows, binary64 exercise decisions, binary32 Greek tangents. Exercise
    decisions agree with ``F64``, so prices agree far inside a standard error.
    ``F32_BULK``: ``FUSED_F64`` with binary32 spots and cashflows. Rounding flips rare near-tie
    exercise decisions and backward induction amplifies them: most strikes stay within 0.01 SE of
    ``F64``, single strikes can move by 0.1-0.5 SE.
    """

    F64 = "f64"
    F32_BULK = "f32_bulk"
    FUSED_F64 = "fused_f64"


_PRECISION_CODE = {HipPrecision.F64: 0, HipPrecision.F32_BULK: 1, HipPrecision.FUSED_F64: 2}


class _Handle:
    """Owns one native executor; destroyed with its Python owner."""

    def __init__(self, device: int, precision: HipPrecision) -> None:
        handle = C.c_void_p()
        err = FohipError()
        code = lib().fohip_executor_create(device, _PRECISION_CODE[precision], C.byref(handle),
                                           C.byref(err))
        raise_for(code, err, lsm=False)
        self.ptr = handle
        self._finalizer = weakref.finalize(self, lib().fohip_executor_destroy, handle)


@dataclass(frozen=True, slots=True)
class HipExecutor:
    """Reusable MI355X resource with an exact device and numerical method.

    Parameters mirror ``CpuExecutor(threads, math_backend)``: ``device`` is the HIP device index and
    ``precision`` the explicit :class:`HipPrecision`. Neither is inferred or substituted. With
    ``autotune`` the batch entry points choose the launch parameters per request shape (see
    ``finoptions_hip.autotune``); results are the same bits either way.
    """

    device: int
    precision: HipPrecision = HipPrecision.F64
    autotune: bool = False
    _handle: _Handle = field(init=False, repr=False, compare=False)
    _tuning: tuple = field(init=False, repr=False, compare=False, default=(0, "auto", "serial", 0, "auto"))

    def __post_init__(self) -> None:
        if type(self.device) is not int or self.device < 0:
            msg = f"device must be a nonnegative plain int; observed {self.device!r}"
            raise InputError(msg)
        if type(self.precision) is not HipPrecision:
            msg = f"precision must be HipPrecision; observed {type(self.precision).__name__}"
            raise InputError(msg)
        if type(self.autotune) is not bool:
            raise InputError(f"autotune must be a plain bool; observed {type(self.autotune).__name__}")
        object.__setattr__(self, "_handle", _Handle(self.device, self.precision))

    @property
    def device_name(self) -> str:
        buf = C.create_string_buffer(256)
        lib().fohip_executor_info(self._handle.ptr, None, None, buf, 256)
        return buf.value.decode()

    def profile(self, enabled: bool) -> None:
        """Enable per-kernel HIP-event timing on this executor's stream (resets the counters)."""
        lib().fohip_reset_profile(self._handle.ptr)
        lib().fohip_set_profiling(self._handle.ptr, 1 if enabled else 0)

    def kernel_stats(self) -> dict[str, tuple[int, float]]:
        """Per-kernel (launches, total milliseconds) since profiling was enabled."""
        n = C.c_size_t(0)
        lib().fohip_get_profile(self._handle.ptr, None, 0, C.byref(n))
        buf = (FohipKernelStat * max(n.value, 1))()
        lib().fohip_get_profile(self._handle.ptr, buf, n.value, C.byref(n))
        return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms))
                for i in range(n.value)}

    def tune(self, *, eval_strike_tile: int = 0, fit_cashflows: str = "auto",
             batch_schedule: str = "serial", batch_mib: int = 0, fit_schedule: str = "auto") -> None:
        """Set launch parameters; results are bitwise invariant to them.

        ``eval_strike_tile``: strikes per evaluation launch column (0: by shape). ``fit_cashflows``
        (fused methods): "value_columns" keeps each strike's cashflow per path in HBM,
        "stopping_rows" keeps one byte per strike and path and rebuilds the cashflow from the
        group's spot matrix (8x less cashflow memory; slower on the shapes measured), "auto"
        keeps value columns. ``batch_schedule``: "serial" runs the memory-bounded batches one
        after another; "overlapped" runs one batch's backward induction beside the next batch's
        simulation and the previous batch's evaluation on a second stream, in two slots of half
        the serial batch memory. ``batch_mib``: paths and cashflows per batch slot in MiB (0: a
        quarter of device memory, an eighth per slot when overlapped). ``fit_schedule`` (fused
        methods): "per_node" launches one fit pass and one reduce per exercise row; "sweep" runs a
        batch's whole backward induction in one persistent launch where every group holds one
        strike (teams of blocks keep their paths on chip, so each spot row is read once, but meet
        in device memory at every row); "auto" is the per-node launches, faster on the shapes
        measured.
        """
        cash_codes = {"auto": 0, "value_columns": 1, "stopping_rows": 2}
        schedule_codes = {"serial": 0, "overlapped": 1}
        fit_codes = {"auto": 0, "per_node": 1, "sweep": 2}
        if fit_cashflows not in cash_codes:
            raise InputError(f"fit_cashflows must be one of {sorted(cash_codes)}; observed {fit_cashflows!r}")
        if batch_schedule not in schedule_codes:
            raise InputError(
                f"batch_schedule must be one of {sorted(schedule_codes)}; observed {batch_schedule!r}")
        if fit_schedule not in fit_codes:
            raise InputError(f"fit_schedule must be one of {sorted(fit_codes)}; observed {fit_schedule!r}")
        if type(batch_mib) is not int or not 0 <= batch_mib < 2**32:
            raise InputError(f"batch_mib must be a plain int in [0, 2^32); observed {batch_mib!r}")
        t = FohipTuning(int(eval_strike_tile), cash_codes[fit_cashflows], schedule_codes[batch_schedule],
                        batch_mib, fit_codes[fit_schedule])
        if lib().fohip_set_tuning(self._handle.ptr, C.byref(t)) != 0:
            raise InputError(f"unsupported tuning eval_strike_tile={eval_strike_tile}")
        object.__setattr__(self, "_tuning",
                           (int(eval_strike_tile), fit_cashflows, batch_schedule, batch_mib, fit_schedule))

    @property
    def tuning(self) -> tuple[int, str, str, int, str]:
        """The launch parameters last set with :meth:`tune` (tile, cashflow encoding, batch
        schedule, slot MiB, fit schedule)."""
        return self._tuning

    def set_near_tie_window(self, window: float = 0.0) -> None:
        """Relative window of the tie counts reported with every result.

        The default 0 counts exact ties, intrinsic == continuation: the decisions the tie rule
        settles and the other rule would decide the other way. With window > 0 a rule application
        counts when |intrinsic - continuation| <= window * |continuation|; 2^-40 (about 9.1e-13)
        is wider than the rounding differences between the methods' continuations, so a decision
        outside it cannot flip between them. That mode is a diagnostic: the evaluation then runs
        single-strike tiles and is several times slower. Prices and every other output are
        independent of the window.
        """
        w = float(window)
        if lib().fohip_set_near_tie_window(self._handle.ptr, C.c_double(w)) != 0:
            raise InputError(f"near-tie window must be finite and in [0, 1]; observed {window!r}")


class CpuExecutor:
    """The reference's CPU resource is not provided by this package.

    Pricing requests name their executor explicitly; this package runs only on HIP devices and
    never substitutes one resource for another.
    """

    def __init__(self, *args: object, **kwargs: object) -> None:
        msg = ("finoptions_hip executes on HIP devices only; use HipExecutor(device=..., "
               "precision=...). No CPU executor or fallback exists in this package.")
        raise ExecutionError(msg)


@dataclass(frozen=True, slots=True)
class RandomStream:
    """A Philox key (``seed``) and counter lane (``domain``) naming one path population: path p
    of the population is the counter stream (block, p, domain) under that key."""

    seed: int
    domain: int

    def __post_init__(self) -> None:
        if type(self.seed) is not int or not 0 <= self.seed <= _MAX_SEED:
            msg = f"seed must be a plain int in [0, {_MAX_SEED}]; observed {self.seed!r}"
            raise InputError(msg)
        if type(self.domain) is not int or not 0 <= self.domain <= _MAX_DOMAIN:
            msg = f"domain must be a plain int in [0, {_MAX_DOMAIN}]; observed {self.domain!r}"
            raise InputError(msg)


@dataclass(frozen=True, slots=True)
class GbmPaths:
    """Simulated states, one row per observation node (exercise nodes, then expiry) and one column
    per path; an exercise row holds the state before that node's cash dividend."""

    observation_nodes: tuple[int, ...]
    states: Float64Matrix

    def __post_init__(self) -> None:
        if (type(self.observation_nodes) is not tuple or not self.observation_nodes
                or any(type(n) is not int or n < 0 or n > _MAX_DOMAIN
                       for n in self.observation_nodes)):
            raise InputError("observation_nodes must be an immutable tuple[int, ...]")
        if any(right <= left for left, right in pairwise(self.observation_nodes)):
            raise InputError("observation_nodes must be strictly increasing")
        if type(self.states) is not np.ndarray or self.states.dtype != np.float64:
            raise InputError("states must be a float64 numpy.ndarray")
        if self.states.ndim != 2 or self.states.shape[0] != len(self.observation_nodes):
            raise InputError(
                f"states must have shape [observation_nodes, paths]; observed {self.states.shape!r}")

    @property
    def path_count(self) -> int:
        return self.states.shape[1]

    def values_at(self, node: int) -> npt.NDArray[np.float64]:
        if type(node) is not int:
            raise InputError("node must be a plain int")
        return self.states[self.observation_nodes.index(node)]


def _path_plan(initial_spot, schedule, curve, dividend_yield_curve, volatility, dividend_policy,
               paths, random_stream, executor):
    """Checked inputs of ``simulate_gbm_seeded`` as a C path plan (python/finoptions/
    simulation.py:262-332 of the reference); returns (plan, arrays kept alive, observation rows)."""
    if type(initial_spot) is not float:
        raise InputError("initial_spot must be a plain float")
    for name, value, cls in (("schedule", schedule, EventSchedule), ("curve", curve, LogDiscountCurve),
                             ("dividend_yield_curve", dividend_yield_curve,
                              CumulativeDividendYieldCurve),
                             ("volatility", volatility, ConstantVolatility),
                             ("dividend_policy", dividend_policy, DividendPolicy),
                             ("random_stream", random_stream, RandomStream),
                             ("executor", executor, HipExecutor)):
        if type(value) is not cls:
            raise InputError(f"{name} must be {cls.__name__}; observed {type(value).__name__}")
    if type(paths) is not int or not 1 <= paths <= _MAX_PATHS:
        raise InputError(f"paths must be a plain int in [1, {_MAX_PATHS}]; observed {paths!r}")
    if schedule.times != curve.times or schedule.times != dividend_yield_curve.times:
        raise SimulationError("discount curve and event schedule must use the identical time grid")
    times = np.ascontiguousarray(schedule.times, dtype=np.float64)
    logd = np.ascontiguousarray(curve.log_discounts, dtype=np.float64)
    q = np.ascontiguousarray(dividend_yield_curve.cumulative_yields, dtype=np.float64)
    dn = np.ascontiguousarray([d.node for d in schedule.dividends] or [0], dtype=np.uint32)
    da = np.ascontiguousarray([d.amount for d in schedule.dividends] or [0.0], dtype=np.float64)
    en = np.ascontiguousarray(list(schedule.exercise_nodes) or [0], dtype=np.uint32)
    plan = FohipPathPlan(ptr(times, C.c_double), times.size, ptr(logd, C.c_double),
                         ptr(q, C.c_double), ptr(dn, C.c_uint32), ptr(da, C.c_double),
                         len(schedule.dividends), ptr(en, C.c_uint32), len(schedule.exercise_nodes),
                         schedule.expiry_node, volatility.annualized_decimal,
                         0 if dividend_policy is DividendPolicy.LIQUIDATOR else 1)
    return plan, (times, logd, q, dn, da, en), len(schedule.exercise_nodes) + 1


def simulate_gbm_seeded(  # noqa: PLR0913 - financial inputs remain independently typed
    *,
    initial_spot: float,
    schedule: EventSchedule,
    curve: LogDiscountCurve,
    dividend_yield_curve: CumulativeDividendYieldCurve,
    volatility: ConstantVolatility,
    dividend_policy: DividendPolicy,
    paths: int,
    random_stream: RandomStream,
    executor: HipExecutor,
) -> GbmPaths:
    """Generate seeded GBM paths on the device and return the time-major observation matrix.

    Mirrors the reference ``simulate_gbm_seeded`` (python/finoptions/simulation.py:262-332).
    """
    plan, _keep, obs = _path_plan(initial_spot, schedule, curve, dividend_yield_curve, volatility,
                                  dividend_policy, paths, random_stream, executor)
    out = np.empty((obs, paths), dtype=np.float64)
    err = FohipError()
    code = lib().fohip_simulate_gbm_seeded(executor._handle.ptr, C.byref(plan), initial_spot,  # noqa: SLF001
                                           paths, random_stream.seed, random_stream.domain,
                                           ptr(out, C.c_double), C.byref(err))
    raise_for(code, err, lsm=False)
    out.flags.writeable = False
    return GbmPaths(observation_nodes=tuple(schedule.exercise_nodes) + (schedule.expiry_node,),
                    states=out)


@dataclass(frozen=True, slots=True)
class DeviceGbmPaths:
    """Time-major pre-event states left in device memory: ``states`` is a torch tensor of shape
    [observation_nodes, paths] on the executor's device."""

    observation_nodes: tuple[int, ...]
    states: object

    @property
    def path_count(self) -> int:
        return int(self.states.shape[1])

    def values_at(self, node: int):
        if type(node) is not int:
            raise InputError("node must be a plain int")
        return self.states[self.observation_nodes.index(node)]


def simulate_gbm_seeded_tensor(  # noqa: PLR0913 - financial inputs remain independently typed
    *,
    initial_spot: float,
    schedule: EventSchedule,
    curve: LogDiscountCurve,
    dividend_yield_curve: CumulativeDividendYieldCurve,
    volatility: ConstantVolatility,
    dividend_policy: DividendPolicy,
    paths: int,
    random_stream: RandomStream,
    executor: HipExecutor,
    dtype: str = "float64",
) -> DeviceGbmPaths:
    """``simulate_gbm_seeded`` with the observation matrix written straight into a torch tensor on
    the executor's device (no host copy), ordered after the work already queued on torch's current
    stream. ``dtype`` "float32" rounds each binary64 state once."""
    import torch

    if dtype not in ("float64", "float32"):
        raise InputError(f"dtype must be 'float64' or 'float32'; observed {dtype!r}")
    plan, _keep, obs = _path_plan(initial_spot, schedule, curve, dividend_yield_curve, volatility,
                                  dividend_policy, paths, random_stream, executor)
    device = torch.device("cuda", executor.device)
    states = torch.empty((obs, paths), dtype=getattr(torch, dtype), device=device)
    stream = torch.cuda.current_stream(device)
    err = FohipError()
    code = lib().fohip_simulate_gbm_seeded_device(
        executor._handle.ptr, C.byref(plan), initial_spot, paths, random_stream.seed,  # noqa: SLF001
        random_stream.domain, C.c_void_p(states.data_ptr()), 0 if dtype == "float64" else 1,
        C.c_void_p(stream.cuda_stream), C.byref(err))
    raise_for(code, err, lsm=False)
    return DeviceGbmPaths(observation_nodes=tuple(schedule.exercise_nodes) + (schedule.expiry_node,),
                          states=states)


def philox_blocks(executor: HipExecutor, *, seed: int, domain: int, path0: int, paths: int,
                  blocks: int) -> np.ndarray:
    """Raw device Philox4x32-10 words, shape [paths, blocks, 4] (parity probe)."""
    out = np.empty((paths, blocks, 4), dtype=np.uint32)
    err = FohipError()
    code = lib().fohip_philox_blocks(executor._handle.ptr, seed, domain, path0, paths, blocks,  # noqa: SLF001
                                     ptr(out, C.c_uint32), C.byref(err))
    raise_for(code, err, lsm=False)
    return out


def gbm_exp(executor: HipExecutor, x: np.ndarray) -> np.ndarray:
    """exp(x) as the device's GBM step computes it (parity probe)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.empty_like(x)
    err = FohipError()
    code = lib().fohip_gbm_exp_probe(executor._handle.ptr, ptr(x, C.c_double), x.size,  # noqa: SLF001
                                     ptr(out, C.c_double), C.byref(err))
    raise_for(code, err, lsm=False)
    return out


def stream_normals(executor: HipExecutor, *, seed: int, domain: int, path0: int, paths: int,
                   count: int) -> tuple[np.ndarray, np.ndarray]:
    """Device ZIGNOR normals and the u64 draws each consumed, shape [paths, count]."""
    out = np.empty((paths, count), dtype=np.float64)
    draws = np.empty((paths, count), dtype=np.uint32)
    err = FohipError()
    code = lib().fohip_stream_normals(executor._handle.ptr, seed, domain, path0, paths, count,  # noqa: SLF001
                                      ptr(out, C.c_double), ptr(draws, C.c_uint32), C.byref(err))
    raise_for(code, err, lsm=False)
    return out, draws


__all__ = [
    "CpuExecutor",
    "DeviceGbmPaths",
    "ExecutionError",
    "GbmPaths",
    "HipExecutor",
    "HipPrecision",
    "RandomStream",
    "SimulationError",
    "gbm_exp",
    "philox_blocks",
    "simulate_gbm_seeded",
    "simulate_gbm_seeded_tensor",
    "stream_normals",
]
