"""The fit sweep on the GPU (opt-in, fit_schedule="sweep"): a batch's whole fused backward
induction in one persistent launch (teams of blocks keeping their paths on chip, one
device-memory gathering per backward step) gives bitwise the results of the per-node launches:
prices, standard errors, Greeks, stopping counts, tie counts, fitted rules and failures, for
binary64 and binary32 storage, beside the generation kernels of other batches, with windowed tie
counts, both tie rules and escrowed dividends. The per-node launches are in turn pinned to the
oracle by the parity tests."""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from conftest import c1_inputs, c2_columns, hip_groups, hip_method
from test_gpu_schedule import _assert_bitwise

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    from finoptions_hip.workloads import ragged_batch

    # 300 contracts of 2^11..2^15 paths (teams of 1 to 6 blocks) and 4..60 exercise dates
    return ragged_batch(300, seed=7, max_exercise=60, min_paths=2048, max_paths=32768)


def _price(fh, ex, w, greeks=(), **kw):
    from finoptions_hip.workloads import ragged_batch_request

    return fh.price_lsm_call_batch(executor=ex, greeks=greeks, **ragged_batch_request(w, seed=7, **kw))


def _fit_kernels(ex):
    return {k: v for k, v in ex.kernel_stats().items() if k.startswith("k_fit")}


@pytest.mark.parametrize("precision", ["FUSED_F64", "F32_BULK"])
def test_sweep_equals_per_node_launches(small_batch, precision):
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=getattr(fh.HipPrecision, precision))
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    ex.tune(fit_schedule="sweep")
    ex.profile(True)
    swept = _price(fh, ex, small_batch, greeks)
    fits = _fit_kernels(ex)
    print(f"{precision} sweep: {fits}", flush=True)
    assert "k_fit_sweep" in fits and "k_fit_bulk_step" not in fits
    ex.tune(fit_schedule="per_node")
    ex.profile(True)
    per_node = _price(fh, ex, small_batch, greeks)
    fits = _fit_kernels(ex)
    print(f"{precision} per node: {fits}", flush=True)
    assert "k_fit_bulk_step" in fits and "k_fit_sweep" not in fits
    ex.profile(False)
    _assert_bitwise(per_node, swept, precision)


def test_sweep_beside_other_batches(small_batch):
    """Split batches: under the overlapped schedule a batch's sweep shares the device with the
    next batch's simulation and the previous batch's evaluation, so its blocks become resident
    as those drain; under the serial schedule one batch follows another in the same slot."""
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    ex.tune(fit_schedule="per_node")
    ref = _price(fh, ex, small_batch, greeks)
    for schedule, mib in (("overlapped", 16), ("serial", 16), ("overlapped", 1)):
        ex.tune(batch_schedule=schedule, batch_mib=mib, fit_schedule="sweep")
        print(f"sweep, {schedule} batches of {mib} MiB", flush=True)
        _assert_bitwise(ref, _price(fh, ex, small_batch, greeks), f"{schedule} {mib}")


def test_sweep_single_call_equals_per_node():
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    method = hip_method(65_536, 65_536, 7, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO)
    ex.tune(fit_schedule="sweep")
    swept = fh.price_lsm_call(**c1_inputs(), executor=ex, method=method, greeks=greeks)
    ex.tune(fit_schedule="per_node")
    per_node = fh.price_lsm_call(**c1_inputs(), executor=ex, method=method, greeks=greeks)
    for f in dataclasses.fields(swept):
        assert repr(getattr(swept, f.name)) == repr(getattr(per_node, f.name)), f.name
    assert swept.regression_coefficients == per_node.regression_coefficients


def test_sweep_windowed_ties_and_tie_rule_equal_per_node(small_batch):
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    ex.set_near_tie_window(2.0**-30)
    ex.tune(fit_schedule="sweep")
    swept = _price(fh, ex, small_batch, (fh.Greek.DELTA,))
    ex.tune(fit_schedule="per_node")
    per_node = _price(fh, ex, small_batch, (fh.Greek.DELTA,))
    _assert_bitwise(per_node, swept, "windowed ties")
    print(f"windowed training ties: {int(np.sum(swept.near_tie_training_decisions))}", flush=True)
    # the other tie rule, on a single-strike group with escrowed cash dividends
    ex.set_near_tie_window(0.0)
    cols = c2_columns(strikes=1)
    cols.pop("yields")
    cols.update(cash_offsets=np.array([0, 2], dtype=np.int64), cash_times=np.array([0.25, 0.75]),
                cash_amounts=np.array([1.5, 2.5]), escrowed=1)
    mesh = fh.UniformExerciseMesh(base_intervals=40, origin_exercise=fh.OriginExercise.EXCLUDED)
    base = hip_method(40_000, 40_000, 3, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    method = dataclasses.replace(base, tie_rule=fh.ExerciseTieRule.GREATER_OR_EQUAL)
    out = []
    for schedule in ("per_node", "sweep"):
        ex.tune(fit_schedule=schedule)
        out.append(fh.price_lsm_call_grouped_batch(groups=hip_groups(cols), mesh=mesh, executor=ex,
                                                   method=method, greeks=(fh.Greek.DELTA, fh.Greek.VEGA)))
    _assert_bitwise(out[0], out[1], "escrowed, greater or equal")


def test_sweep_failure_equals_per_node(small_batch):
    """With the FAIL support policy some contracts fail in the fit; the sweep ends their teams at
    the failing step and reports the same first failure as the per-node launches."""
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    fail = fh.InsufficientExerciseSupportPolicy.FAIL
    errors = []
    for schedule in ("per_node", "sweep"):
        ex.tune(fit_schedule=schedule)
        with pytest.raises(fh.LsmError) as exc:
            _price(fh, ex, small_batch, support=fail)
        e = exc.value
        errors.append((e.code, e.row, e.node, e.index, str(e)))
    print(f"first failure: {errors[0]}", flush=True)
    assert errors[0] == errors[1]
    # the executor prices normally afterwards
    ex.tune(fit_schedule="per_node")
    ok = _price(fh, ex, small_batch)
    ex.tune(fit_schedule="sweep")
    _assert_bitwise(ok, _price(fh, ex, small_batch), "after failures")


def test_sweep_equals_per_node_at_ragged_batch_team_sizes():
    """600 contracts drawn like the ragged batch's (16 384 to 262 144 paths: teams of 3 to 43
    blocks, 4 to 252 backward steps), with delta and vega."""
    import finoptions_hip as fh
    from finoptions_hip.workloads import ragged_batch, ragged_batch_request

    w = ragged_batch(600, seed=31)
    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    out = []
    for schedule in ("per_node", "sweep"):
        ex.tune(fit_schedule=schedule)
        out.append(fh.price_lsm_call_batch(executor=ex, greeks=greeks, **ragged_batch_request(w, seed=31)))
    print(f"{len(w['paths'])} contracts, up to {int(w['paths'].max())} paths", flush=True)
    _assert_bitwise(out[0], out[1], "ragged team sizes")
