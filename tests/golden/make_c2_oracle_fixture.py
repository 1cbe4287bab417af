"""Writes tests/golden/c2_full_oracle.json: the oracle's FP64 result columns for the full-size
grouped-strike workload (1 state x 64 strikes linspace(80, 120), S0 100, r 0.03, q 0.02,
sigma 0.25, T 1, UniformExerciseMesh(50, EXCLUDED), 2^20 training + 2^20 evaluation paths,
streams (11, 1) / (11, 2), CONTINUE support), so that GPU tests can check that size against the
oracle without re-running it (about two minutes on one CPU thread).

Floating-point columns are stored as hexadecimal binary64 strings (float.hex), exactly.

Usage: python tests/golden/make_c2_oracle_fixture.py
"""

from __future__ import annotations

import json
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "oracle"))

import pyoracle  # noqa: E402

PATHS = 1 << 20
STRIKES = 64


def request() -> pyoracle.GroupedRequest:
    return pyoracle.GroupedRequest(
        group_offsets=np.array([0, STRIKES], dtype=np.uint32), initial_spots=np.array([100.0]),
        strikes=np.linspace(80.0, 120.0, STRIKES), maturities=np.array([1.0]),
        rates=np.array([0.03]), volatilities=np.array([0.25]), dividend_yields=np.array([0.02]),
        training_paths=PATHS, evaluation_paths=PATHS, training_seed=11, evaluation_seed=11,
        training_domain_start=1, evaluation_domain_start=2, mesh_kind=0, base_intervals=50,
        origin_exercise=0, insufficient_support=1)


def main() -> None:
    t0 = time.perf_counter()
    out = pyoracle.price_grouped(request(), threads=1)
    seconds = time.perf_counter() - t0
    fixture = {"workload": "C2 grouped strikes, 64 x 50 intervals, 2^20 + 2^20 paths, seed 11",
               "paths": PATHS, "strikes": STRIKES, "oracle_seconds_one_thread": round(seconds, 1)}
    for name, values in out.items():
        arr = np.asarray(values)
        if arr.dtype == np.float64:
            fixture[name] = [float(v).hex() for v in arr]
        else:
            fixture[name] = [int(v) for v in arr]
    path = ROOT / "tests" / "golden" / "c2_full_oracle.json"
    path.write_text(json.dumps(fixture, indent=0) + "\n")
    print(f"wrote {path} ({seconds:.1f} s)")


if __name__ == "__main__":
    main()
