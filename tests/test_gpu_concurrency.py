"""Concurrent requests on distinct executors of one device: each executor owns its streams and
buffers, so two threads pricing at once get bitwise the results of pricing one after the other.
(ctypes releases the GIL for the library call, so the two requests overlap on the host and the
device.)"""

from __future__ import annotations

import threading

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method

pytestmark = pytest.mark.gpu


def _grouped(fh, ex):
    mesh = fh.UniformExerciseMesh(base_intervals=30, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(65_536, 65_536, 21, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return fh.price_lsm_call_grouped_batch(groups=hip_groups(c2_columns(strikes=32)), mesh=mesh,
                                           executor=ex, method=method, greeks=(fh.Greek.DELTA,))


def _ragged(fh, ex):
    from finoptions_hip.workloads import ragged_batch, ragged_batch_request

    w = ragged_batch(200, seed=9, max_exercise=40, min_paths=2048, max_paths=16384)
    return fh.price_lsm_call_batch(executor=ex, greeks=(fh.Greek.VEGA,), **ragged_batch_request(w, seed=9))


def _bits(r):
    g = next(iter(r.greeks.values()))
    return (r.prices.tobytes(), r.standard_errors.tobytes(), r.early_exercise_paths.tobytes(), g.tobytes())


@pytest.mark.parametrize("precisions", [("FUSED_F64", "FUSED_F64"), ("F64", "FUSED_F64")])
def test_two_executors_price_concurrently_like_sequentially(precisions):
    import finoptions_hip as fh

    a = fh.HipExecutor(device=0, precision=getattr(fh.HipPrecision, precisions[0]))
    b = fh.HipExecutor(device=0, precision=getattr(fh.HipPrecision, precisions[1]))
    ref_a, ref_b = _bits(_grouped(fh, a)), _bits(_ragged(fh, b))
    results, errors = {}, []

    def work(name, fn, ex):
        try:
            results[name] = [_bits(fn(fh, ex)) for _ in range(3)]
        except Exception as exc:   # reported by the main thread
            errors.append(exc)

    threads = [threading.Thread(target=work, args=("a", _grouped, a)),
               threading.Thread(target=work, args=("b", _ragged, b))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not errors, errors
    assert all(r == ref_a for r in results["a"])
    assert all(r == ref_b for r in results["b"])
