"""The allocation-free single-call entry point (fohip_price_lsm_call_noalloc) against the
allocating one: bitwise the same results on two calls in a row, the second smaller, for the
binary64 and the fused method, with Greeks; an oversized request and profiling are refused."""

from __future__ import annotations

import pytest

from conftest import c1_inputs

pytestmark = pytest.mark.gpu


def _method(fh, paths, seeds):
    return fh.Lsm(training_paths=paths, evaluation_paths=paths,
                  training_stream=fh.RandomStream(seed=seeds[0], domain=1),
                  evaluation_stream=fh.RandomStream(seed=seeds[1], domain=2),
                  insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL)


def _same(a, b):
    assert a.price.hex() == b.price.hex()
    assert a.standard_error.hex() == b.standard_error.hex()
    assert a.sample_standard_deviation.hex() == b.sample_standard_deviation.hex()
    assert a.exercise_counts == b.exercise_counts
    assert a.fit_diagnostics == b.fit_diagnostics
    assert a.regression_coefficients == b.regression_coefficients
    assert [(g, v.hex(), s.hex()) for g, v, s in a.greeks] == [(g, v.hex(), s.hex()) for g, v, s in b.greeks]
    assert a.near_tie_training_decisions == b.near_tie_training_decisions


@pytest.mark.parametrize("precision", ["F64", "FUSED_F64"])
def test_allocation_free_call_equals_the_allocating_call(precision):
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=getattr(fh.HipPrecision, precision))
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)
    for paths, seeds in ((65_536, (7, 7)), (12_288, (3, 5))):   # the second call is smaller
        kw = dict(c1_inputs(), executor=ex, method=_method(fh, paths, seeds), greeks=greeks)
        free = fh.price_lsm_call(**kw, allocation_free=True)
        ref = fh.price_lsm_call(**kw)
        print(f"{precision} {paths} paths: allocation-free {free.price!r}, allocating {ref.price!r}")
        _same(free, ref)
        again = fh.price_lsm_call(**kw, allocation_free=True)   # the region keeps the last contents
        _same(again, ref)


def test_allocation_free_call_refuses_what_its_region_cannot_hold():
    import finoptions_hip as fh

    ex = fh.HipExecutor(device=0, precision=fh.HipPrecision.F64)
    kw = dict(c1_inputs(), executor=ex, method=_method(fh, 4_000_000, (7, 7)))
    with pytest.raises(fh.ExecutionError, match="allocation-free entry point holds"):
        fh.price_lsm_call(**kw, allocation_free=True)
    ex.profile(True)
    try:
        with pytest.raises(fh.ExecutionError, match="profiling off"):
            fh.price_lsm_call(**dict(kw, method=_method(fh, 1024, (7, 7))), allocation_free=True)
    finally:
        ex.profile(False)
    # an invalid request fails exactly like the allocating entry point
    bad = dict(kw, method=fh.Lsm(training_paths=64, evaluation_paths=64,
                                  training_stream=fh.RandomStream(seed=7, domain=1),
                                  evaluation_stream=fh.RandomStream(seed=8, domain=2),
                                  insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL,
                                  condition_limit=1.0))
    with pytest.raises(fh.LsmError) as e1:
        fh.price_lsm_call(**bad, allocation_free=True)
    with pytest.raises(fh.LsmError) as e2:
        fh.price_lsm_call(**bad)
    assert str(e1.value) == str(e2.value)
