"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a sanitized driver
(oracle/asan_driver.c) prices small requests of every dividend model and mesh kind, serial and
group-parallel, and exports paths and pathwise Greeks; any sanitizer report fails the run."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("make") is None, reason="needs gcc and make")
def test_oracle_runs_clean_under_sanitizers():
    run = subprocess.run(["make", "-s", "-C", str(ROOT / "oracle"), "asan-check"], capture_output=True,
                         text=True, timeout=600, env={"PATH": "/usr/bin:/bin", "ASAN_OPTIONS": "detect_leaks=1"})
    out = run.stdout + run.stderr
    print(out[-2000:])
    assert run.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out
    assert "sanitized oracle run: 0 failed checks" in out
