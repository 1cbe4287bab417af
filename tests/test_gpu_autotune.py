"""The autotuner on the GPU: an executor created with ``autotune=True`` times its candidates on a
sample of the first request of each shape, installs one, prices bitwise like an untuned executor,
and reuses the cached choice for the next request of the same shape."""

from __future__ import annotations

import numpy as np
import pytest

from conftest import c2_columns, hip_groups, hip_method

pytestmark = pytest.mark.gpu


def _same(a, b):
    assert np.array_equal(a.prices.view(np.uint64), b.prices.view(np.uint64))
    assert np.array_equal(a.standard_errors.view(np.uint64), b.standard_errors.view(np.uint64))
    assert np.array_equal(a.early_exercise_paths, b.early_exercise_paths)
    for g in (a.greeks or {}):
        assert np.array_equal(a.greeks[g].view(np.uint64), b.greeks[g].view(np.uint64))


def test_grouped_request_is_tuned_once_and_prices_like_the_default():
    import finoptions_hip as fh
    from finoptions_hip import autotune

    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    mesh = fh.UniformExerciseMesh(base_intervals=50, origin_exercise=fh.OriginExercise.EXCLUDED)
    method = hip_method(65_536, 65_536, 11, support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    groups = hip_groups(c2_columns(strikes=64))
    plain = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    tuned = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64, autotune=True)
    ref = fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=plain, method=method,
                                          greeks=greeks)
    before = autotune.cache().tuned
    got = fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=tuned, method=method,
                                          greeks=greeks)
    assert autotune.cache().tuned == before + 1
    key = autotune.shape_key(tuned, groups.group_offsets, mesh, groups.dividends, method, greeks, None)
    rec = autotune.cache().records[key]
    print("grouped choice", rec.choice, {k: round(v * 1e3, 2) for k, v in rec.timings.items()}, flush=True)
    assert rec.choice["eval_strike_tile"] in autotune.GREEK_TILES
    assert rec.choice["fit_cashflows"] in ("value_columns", "stopping_rows")
    assert len(rec.timings) >= len(autotune.GREEK_TILES)
    assert tuned.tuning[:2] == (rec.choice["eval_strike_tile"], rec.choice["fit_cashflows"])
    _same(ref, got)
    again = fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=tuned, method=method,
                                            greeks=greeks)
    assert autotune.cache().tuned == before + 1        # cache hit: no second sample
    _same(ref, again)


def test_ragged_batch_is_tuned_and_prices_like_the_default():
    import finoptions_hip as fh
    from finoptions_hip import autotune
    from finoptions_hip.workloads import ragged_batch, ragged_batch_request

    w = ragged_batch(300, seed=7, max_exercise=60, min_paths=2048, max_paths=32768)
    req = ragged_batch_request(w, seed=7)
    greeks = (fh.Greek.DELTA,)
    plain = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    tuned = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64, autotune=True)
    ref = fh.price_lsm_call_batch(executor=plain, greeks=greeks, **req)
    before = autotune.cache().tuned
    got = fh.price_lsm_call_batch(executor=tuned, greeks=greeks, **req)
    assert autotune.cache().tuned == before + 1
    _same(ref, got)
    # single-strike groups: the tile candidates are the engine's choice and tile 1
    key = autotune.shape_key(tuned, np.arange(301, dtype=np.uint32), req["mesh"], req["dividends"],
                             req["method"], greeks, req["path_counts"])
    rec = autotune.cache().records[key]
    print("ragged choice", rec.choice, rec.sample_groups, flush=True)
    assert {t for t, _ in rec.timings} == {0, 1}
    assert 1 <= rec.sample_groups < 300
