"""C-ABI boundary checks that need no GPU: symbols, validation and fail-fast behaviour."""

from __future__ import annotations

import ctypes as C
import re

import numpy as np
import pytest

from conftest import ROOT


def header_symbols() -> set[str]:
    text = (ROOT / "include" / "fohip.h").read_text()
    return set(re.findall(r"^\s*(?:int|void)\s+(fohip_\w+)\(", text, re.M))


def test_library_exports_every_declared_symbol():
    import finoptions_hip._lib as L

    handle = C.CDLL(str(L.LIB_PATH))
    declared = header_symbols()
    assert declared == set(L.EXPORTED)
    for name in sorted(declared):
        assert hasattr(handle, name), name
    assert handle.fohip_abi_version() == 5


def test_library_has_gfx950_code_object():
    data = (ROOT / "finoptions-engine_amd" / "finoptions_hip" / "libfohip.so").read_bytes()
    assert b"gfx950" in data


def test_executor_without_device_fails_fast(monkeypatch):
    import finoptions_hip as fh

    if fh.device_count() > 0:
        pytest.skip("a HIP device is visible")
    with pytest.raises(fh.ExecutionError, match="no HIP device"):
        fh.HipExecutor(device=0)
    with pytest.raises(fh.ExecutionError, match="HIP devices only"):
        fh.CpuExecutor(threads=1)


def test_financial_validation_mirrors_the_reference():
    import finoptions_hip as fh

    with pytest.raises(fh.InputError, match="strictly increasing"):
        fh.EventSchedule(times=(0.0, 0.5, 0.5), dividends=(), exercise_nodes=(), expiry_node=2)
    with pytest.raises(fh.InputError, match="canonical 0.0"):
        fh.LogDiscountCurve(times=(0.0, 1.0), log_discounts=(-0.0, -0.01))
    with pytest.raises(fh.InputError):
        fh.CumulativeDividendYieldCurve(times=(0.0, 1.0), cumulative_yields=(0.0, -0.01))
    with pytest.raises(fh.InputError):
        fh.ConstantVolatility(0.0)
    with pytest.raises(fh.InputError):
        fh.CashDividend(node=1, amount=-1.0)
    with pytest.raises(fh.InputError, match="precede expiry"):
        fh.EventSchedule(times=(0.0, 1.0), dividends=(), exercise_nodes=(1,), expiry_node=1)
    with pytest.raises(fh.InputError, match="duplicate cash dividend"):
        fh.EventSchedule(times=(0.0, 0.5, 1.0),
                         dividends=(fh.CashDividend(node=1, amount=1.0),
                                    fh.CashDividend(node=1, amount=2.0)),
                         exercise_nodes=(), expiry_node=2)
    ok = fh.EventSchedule(times=(0.0, 0.5, 1.0), dividends=(fh.CashDividend(node=2, amount=1.0),),
                          exercise_nodes=(0, 1), expiry_node=2)
    assert ok.expiry_node == 2


def test_request_types_reject_inference():
    import finoptions_hip as fh

    with pytest.raises(fh.InputError):
        fh.RandomStream(seed=-1, domain=0)
    with pytest.raises(fh.InputError):
        fh.Lsm(training_paths=10, evaluation_paths=10,
               training_stream=fh.RandomStream(seed=1, domain=1),
               evaluation_stream=fh.RandomStream(seed=1, domain=1),
               insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL)
    with pytest.raises(fh.InputError):
        fh.Lsm(training_paths=10, evaluation_paths=10,
               training_stream=fh.RandomStream(seed=1, domain=1),
               evaluation_stream=fh.RandomStream(seed=1, domain=2),
               insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL,
               polynomial_degree=5)
    with pytest.raises(fh.InputError, match="dtype"):
        fh.LsmStrikeGroups(group_offsets=np.array([0, 1], dtype=np.int64),
                           initial_spots=np.array([1.0]), strikes=np.array([1.0]),
                           maturities=np.array([1.0]), rates=np.array([0.0]),
                           volatilities=np.array([0.2]),
                           dividends=fh.ContinuousDividendYieldBatch(
                               annualized_decimals=np.array([0.0])))
    with pytest.raises(fh.InputError, match="C-contiguous"):
        fh.ContinuousDividendYieldBatch(annualized_decimals=np.zeros(4)[::2])


def test_ragged_offsets_accept_sliced_arrow_lists():
    import finoptions_hip as fh

    zero = fh.RaggedCashDividends(offsets=np.array([0, 2, 2, 3], dtype=np.int64),
                                  times=np.array([0.1, 0.2, 0.4]),
                                  amounts=np.array([1.0, 2.0, 3.0]))
    sliced = fh.RaggedCashDividends(offsets=np.array([7, 9, 9, 10], dtype=np.int64),
                                    times=np.array([0.1, 0.2, 0.4]),
                                    amounts=np.array([1.0, 2.0, 3.0]))
    assert zero.rows == sliced.rows == 3
    with pytest.raises(fh.InputError):
        fh.RaggedCashDividends(offsets=np.array([0, 2, 1], dtype=np.int64),
                               times=np.array([0.1, 0.2]), amounts=np.array([1.0, 2.0]))


def test_product_does_not_import_the_oracle():
    pkg = ROOT / "finoptions-engine_amd"
    for path in list(pkg.rglob("*.py")) + list(pkg.rglob("*.cpp")) + list(pkg.rglob("*.hip")):
        text = path.read_text()
        assert "pyoracle" not in text and "lsm_oracle" not in text, path


STRUCTS = {
    "fohip_error": "FohipError", "fohip_tuning": "FohipTuning", "fohip_kernel_stat": "FohipKernelStat",
    "fohip_call_request": "FohipCallRequest", "fohip_node_diagnostic": "FohipNodeDiagnostic",
    "fohip_call_result": "FohipCallResult", "fohip_grouped_request": "FohipGroupedRequest",
    "fohip_batch_result": "FohipBatchResult", "fohip_path_plan": "FohipPathPlan",
}


def test_ctypes_layouts_match_the_header(tmp_path):
    """Every ctypes mirror has the C struct's size and each field's offset, as gcc lays them out
    from include/fohip.h (the field names are the header's)."""
    import shutil
    import subprocess

    import finoptions_hip._lib as L

    if shutil.which("gcc") is None:
        pytest.skip("gcc is not available")
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "fohip.h"', "int main(void) {"]
    for cname, pyname in STRUCTS.items():
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for field, *_ in getattr(L, pyname)._fields_:
            lines.append(f'    printf("{cname} {field} %zu\\n", offsetof({cname}, {field}));')
    lines += ["    return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", f"-I{ROOT / 'include'}", str(src), "-o", str(exe)], check=True,
                   capture_output=True, text=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    c_layout = {tuple(line.split()[:2]): int(line.split()[2]) for line in out if line}
    for cname, pyname in STRUCTS.items():
        cls = getattr(L, pyname)
        assert C.sizeof(cls) == c_layout[(cname, "size")], cname
        for field, *_ in cls._fields_:
            assert getattr(cls, field).offset == c_layout[(cname, field)], (cname, field)


@pytest.mark.parametrize("base", [0, 5])
def test_ragged_schedules_report_the_first_failing_row_and_check(base):
    """Per-row checks of RaggedExerciseTimes and RaggedCashDividends (python/finoptions/
    batch_lsm.py, dividends.py): within a row the time checks come before the amount checks, and
    the first failing row in row order is named; sliced offsets address the same rows."""
    import finoptions_hip as fh

    off = np.array([0, 2, 2, 4, 6], dtype=np.int64) + base

    def ex(times):
        return fh.RaggedExerciseTimes(offsets=off, times=np.array(times))

    def cash(times, amounts):
        return fh.RaggedCashDividends(offsets=off, times=np.array(times), amounts=np.array(amounts))

    ex([0.1, 0.2, 0.3, 0.4, 0.5, 0.6])
    # offsets [0, 2, 2, 4, 6]: row 1 is empty, row 2 holds times 2..3 and row 3 times 4..5
    with pytest.raises(fh.InputError, match=r"^row 3 exercise times must be finite and positive$"):
        ex([0.1, 0.2, 0.3, 0.4, -1.0, 0.6])
    with pytest.raises(fh.InputError, match=r"^row 2 exercise times must be strictly increasing$"):
        ex([0.1, 0.2, 0.3, 0.3, 0.5, 0.6])
    # row 2 is increasing but row 3 holds a NaN: the earlier row's check passes first
    with pytest.raises(fh.InputError, match=r"^row 3 exercise times must be finite and positive$"):
        ex([0.1, 0.2, 0.3, 0.4, 0.5, float("nan")])
    # a row failing both checks reports the positivity check
    with pytest.raises(fh.InputError, match=r"^row 0 exercise times must be finite and positive$"):
        ex([0.2, -0.1, 0.3, 0.4, 0.5, 0.6])

    cash([0.1, 0.1, 0.3, 0.4, 0.5, 0.6], [1.0, 0.0, 1.0, 1.0, 1.0, 1.0])   # ties and zero are fine
    with pytest.raises(fh.InputError, match=r"^row 2 cash-dividend times must be nondecreasing$"):
        cash([0.1, 0.2, 0.4, 0.3, 0.5, 0.6], [1.0] * 6)
    with pytest.raises(fh.InputError, match=r"^row 0 cash-dividend times must be finite and positive$"):
        cash([0.0, 0.2, 0.3, 0.4, 0.5, 0.6], [1.0] * 6)
    with pytest.raises(fh.InputError, match=r"^row 3 cash-dividend amounts must be finite and nonnegative$"):
        cash([0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [1.0, 1.0, 1.0, 1.0, 1.0, -0.0])
    with pytest.raises(fh.InputError, match=r"^row 2 cash-dividend amounts must be finite and nonnegative$"):
        cash([0.1, 0.2, 0.3, 0.4, 0.5, 0.6], [1.0, 1.0, float("inf"), 1.0, 1.0, 1.0])
    # the time check of a row precedes its amount check
    with pytest.raises(fh.InputError, match=r"^row 2 cash-dividend times must be nondecreasing$"):
        cash([0.1, 0.2, 0.4, 0.3, 0.5, 0.6], [1.0, 1.0, float("nan"), 1.0, 1.0, 1.0])
