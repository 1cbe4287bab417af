"""Shared pytest configuration: import paths, the `gpu` marker, executors and request builders."""

from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT / "finoptions-engine_amd"))
sys.path.insert(0, str(ROOT / "oracle"))

GOLDEN = json.loads((ROOT / "tests" / "golden" / "reference_vectors.json").read_text())


def pytest_configure(config: pytest.Config) -> None:
    config.addinivalue_line("markers", "gpu: needs a HIP (MI355X) device")
    lib = ROOT / "oracle" / "lib" / "liblsm_oracle.so"
    if not lib.exists():
        subprocess.run(["make", "-s", "-C", str(ROOT / "oracle")], check=True)


@pytest.fixture(scope="session")
def oracle():
    import pyoracle

    return pyoracle


@pytest.fixture(scope="session")
def golden() -> dict:
    return GOLDEN


@pytest.fixture(scope="session")
def hip_f64():
    import finoptions_hip as fh

    return fh.HipExecutor(device=0, precision=fh.HipPrecision.F64)


@pytest.fixture(scope="session")
def hip_f32():
    import finoptions_hip as fh

    return fh.HipExecutor(device=0, precision=fh.HipPrecision.F32_BULK)


@pytest.fixture(scope="session")
def hip_fused():
    import finoptions_hip as fh

    return fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)


# ------------------------------------------------------------------ canonical request builders

def c1_inputs():
    """The single scheduled call of the headline CPU case: S0 = K = 100, times (0, 0.5, 1), one
    cash dividend 2.0 at node 1 (liquidator), exercise node 1, log-discounts (0, -0.01, -0.025),
    zero yield, sigma 0.20."""
    import finoptions_hip as fh

    schedule = fh.EventSchedule(times=(0.0, 0.5, 1.0),
                                dividends=(fh.CashDividend(node=1, amount=2.0),),
                                exercise_nodes=(1,), expiry_node=2)
    curve = fh.LogDiscountCurve(times=schedule.times, log_discounts=(0.0, -0.01, -0.025))
    return {
        "initial_spot": 100.0,
        "strike": 100.0,
        "schedule": schedule,
        "curve": curve,
        "dividend_yield_curve": fh.CumulativeDividendYieldCurve.zero(times=schedule.times),
        "volatility": fh.ConstantVolatility(0.20),
        "dividend_policy": fh.DividendPolicy.LIQUIDATOR,
    }


def c1_plan():
    import pyoracle

    return pyoracle.Plan(times=[0.0, 0.5, 1.0], log_discounts=[0.0, -0.01, -0.025],
                         cum_yields=[0.0, 0.0, 0.0], volatility=0.20, dividend_policy=0,
                         dividend_nodes=[1], dividend_amounts=[2.0], exercise_nodes=[1],
                         expiry_node=2)


def c2_columns(strikes: int = 64, spot: float = 100.0):
    """Grouped-strike state: S0 100, r 0.03, q 0.02, sigma 0.25, T 1, strikes linspace(80, 120)."""
    return {
        "group_offsets": np.array([0, strikes], dtype=np.uint32),
        "initial_spots": np.array([spot]),
        "strikes": np.linspace(80.0, 120.0, strikes),
        "maturities": np.array([1.0]),
        "rates": np.array([0.03]),
        "volatilities": np.array([0.25]),
        "yields": np.array([0.02]),
    }


def oracle_grouped(cols: dict, *, base_intervals: int, training_paths: int,
                   evaluation_paths: int, seed: int, train_domain: int = 1,
                   eval_domain: int = 2, origin: int = 0, **kw):
    import pyoracle

    return pyoracle.GroupedRequest(
        group_offsets=cols["group_offsets"], initial_spots=cols["initial_spots"],
        strikes=cols["strikes"], maturities=cols["maturities"], rates=cols["rates"],
        volatilities=cols["volatilities"], dividend_yields=cols.get("yields"),
        cash_offsets=cols.get("cash_offsets"), cash_times=cols.get("cash_times"),
        cash_amounts=cols.get("cash_amounts"), cash_policy=cols.get("cash_policy", 0),
        escrowed=cols.get("escrowed"), training_paths=training_paths, evaluation_paths=evaluation_paths, training_seed=seed,
        evaluation_seed=seed, training_domain_start=train_domain,
        evaluation_domain_start=eval_domain, mesh_kind=kw.pop("mesh_kind", 0),
        base_intervals=base_intervals, origin_exercise=kw.pop("origin_exercise", origin), **kw)


def hip_groups(cols: dict):
    import finoptions_hip as fh

    if "cash_offsets" in cols and cols.get("escrowed") is not None:
        dividends = fh.EscrowedCashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=cols["cash_offsets"], times=cols["cash_times"],
                                             amounts=cols["cash_amounts"]),
            volatility_convention=fh.EscrowedVolatilityConvention.PREPAID_FORWARD
            if cols["escrowed"] == 0 else fh.EscrowedVolatilityConvention.SPOT_MATCHED_AT_ORIGIN,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    elif "cash_offsets" in cols:
        dividends = fh.CashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=cols["cash_offsets"], times=cols["cash_times"],
                                             amounts=cols["cash_amounts"]),
            policy=fh.DividendPolicy.LIQUIDATOR if cols.get("cash_policy", 0) == 0
            else fh.DividendPolicy.SURVIVOR,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE)
    else:
        dividends = fh.ContinuousDividendYieldBatch(annualized_decimals=cols["yields"])
    return fh.LsmStrikeGroups(group_offsets=cols["group_offsets"],
                              initial_spots=cols["initial_spots"], strikes=cols["strikes"],
                              maturities=cols["maturities"], rates=cols["rates"],
                              volatilities=cols["volatilities"], dividends=dividends)


def hip_method(training_paths: int, evaluation_paths: int, seed: int, train_domain: int = 1,
               eval_domain: int = 2, support=None, degree: int = 4):
    import finoptions_hip as fh

    return fh.Lsm(training_paths=training_paths, evaluation_paths=evaluation_paths,
                  training_stream=fh.RandomStream(seed=seed, domain=train_domain),
                  evaluation_stream=fh.RandomStream(seed=seed, domain=eval_domain),
                  insufficient_exercise_support=support
                  or fh.InsufficientExerciseSupportPolicy.FAIL,
                  polynomial_degree=degree)


def rel_err(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    denom = np.maximum(np.abs(b), 1e-300)
    return float(np.max(np.abs(a - b) / denom)) if a.size else 0.0
