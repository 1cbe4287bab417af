"""Batch schedules on the GPU: a request split into many memory-bounded batches gives bitwise the
results of one batch, whether the batches overlap on two streams or run one after another, and a
failure in a late batch is reported as the first failure in request order under every schedule.
"""

from __future__ import annotations

import dataclasses

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

# small ragged batch: 300 contracts of 2^11..2^15 paths and 4..60 exercise dates, about 3 MB of
# paths per contract, so a 16 MiB slot holds a handful of contracts and a 1 MiB slot one
CONTRACTS = 300


@pytest.fixture(scope="module")
def small_batch():
    from finoptions_hip.workloads import ragged_batch

    return ragged_batch(CONTRACTS, seed=7, max_exercise=60, min_paths=2048, max_paths=32768)


def _price(fh, ex, w, greeks=(), **kw):
    from finoptions_hip.workloads import ragged_batch_request

    return fh.price_lsm_call_batch(executor=ex, greeks=greeks, **ragged_batch_request(w, seed=7, **kw))


def _assert_bitwise(a, b, label):
    for f in dataclasses.fields(a):
        x, y = getattr(a, f.name), getattr(b, f.name)
        if isinstance(x, dict):
            assert x.keys() == y.keys(), (label, f.name)
            for k in x:
                assert np.array_equal(x[k].view(np.uint64), y[k].view(np.uint64)), (label, f.name, k)
        elif x is None:
            assert y is None, (label, f.name)
        elif np.asarray(x).dtype == np.float64:
            assert np.array_equal(np.asarray(x).view(np.uint64), np.asarray(y).view(np.uint64)), (label, f.name)
        else:
            assert np.array_equal(x, y), (label, f.name)


@pytest.mark.parametrize("precision", ["FUSED_F64", "F64"])
def test_split_batches_equal_one_batch_under_both_schedules(small_batch, precision):
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=getattr(fh.HipPrecision, precision))
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    one = _price(fh, ex, small_batch, greeks)   # default slot: the whole request in one batch
    for schedule, mib in (("overlapped", 16), ("serial", 16), ("overlapped", 1)):
        ex.tune(batch_schedule=schedule, batch_mib=mib)
        split = _price(fh, ex, small_batch, greeks)
        print(f"{precision} {schedule} {mib} MiB: bitwise check", flush=True)
        _assert_bitwise(one, split, f"{schedule} {mib} MiB")
    # the executor is reusable after split requests, in either order of schedules
    ex.tune()
    _assert_bitwise(one, _price(fh, ex, small_batch, greeks), "default after split")


def _oracle_request(oracle, w):
    """The whole small batch as one oracle row batch with the FAIL support policy (group g keeps
    the stream domains of row g: training 1 + g, evaluation 2 + rows + g)."""
    rows = len(w["initial_spots"])
    return oracle.GroupedRequest(
        group_offsets=np.arange(rows + 1, dtype=np.uint32), initial_spots=w["initial_spots"],
        strikes=w["strikes"], maturities=w["maturities"], rates=w["rates"],
        volatilities=w["volatilities"], cash_offsets=w["cash_offsets"], cash_times=w["cash_times"],
        cash_amounts=w["cash_amounts"], cash_policy=0, training_paths=int(w["paths"].max()),
        evaluation_paths=int(w["paths"].max()), training_path_counts=w["paths"],
        evaluation_path_counts=w["paths"], training_seed=7, evaluation_seed=7,
        training_domain_start=1, evaluation_domain_start=2 + rows, mesh_kind=2,
        exercise_offsets=w["exercise_offsets"], exercise_times=w["exercise_times"], origin_exercise=0,
        insufficient_support=0)


def test_late_batch_failure_is_the_first_failure_under_every_schedule(small_batch, oracle):
    """With the FAIL support policy some contracts fail in the fit (row 11 is the first, in the
    oracle's processing order); the error names the first failing row in request order, which
    lies in a late batch when slots are small."""
    import finoptions_hip as fh

    with pytest.raises(oracle.OracleError) as cpu:
        oracle.price_grouped(_oracle_request(oracle, small_batch))
    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    fail = fh.InsufficientExerciseSupportPolicy.FAIL
    with pytest.raises(fh.LsmError) as whole:
        _price(fh, ex, small_batch, support=fail)
    ref, c = whole.value, cpu.value
    print(f"first failure: code {ref.code} row {ref.row} node {ref.node}; oracle {c.code} {c.row} {c.node}",
          flush=True)
    assert (ref.code, ref.row, ref.node, ref.index) == (c.code, c.row, c.node, c.index)
    assert ref.row > 0
    for schedule, mib in (("overlapped", 16), ("serial", 16), ("overlapped", 1)):
        ex.tune(batch_schedule=schedule, batch_mib=mib)
        with pytest.raises(fh.LsmError) as split:
            _price(fh, ex, small_batch, support=fail)
        e = split.value
        assert (e.code, e.row, e.node, e.index, str(e)) == (ref.code, ref.row, ref.node, ref.index, str(ref))
    # and a request that succeeds still does after the failed ones
    ex.tune(batch_mib=16)
    ok = _price(fh, ex, small_batch)
    ex.tune()
    _assert_bitwise(ok, _price(fh, ex, small_batch), "after failures")
