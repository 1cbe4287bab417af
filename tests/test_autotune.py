"""Autotuner host logic (no GPU): the sample is the request's first groups with their own stream
domains, sized by path work, and shapes key the cache by kind and rounded size."""

from __future__ import annotations

import numpy as np


def _columns(req):
    rows = req["initial_spots"].size
    return dict(group_offsets=np.arange(rows + 1, dtype=np.uint32), initial_spots=req["initial_spots"],
                strikes=req["strikes"], maturities=req["maturities"], rates=req["rates"],
                volatilities=req["volatilities"], dividends=req["dividends"], mesh=req["mesh"],
                path_counts=req["path_counts"])


def test_sample_is_the_request_built_for_its_first_contracts():
    from finoptions_hip import autotune
    from finoptions_hip.workloads import ragged_batch, ragged_batch_request

    w = ragged_batch(50, seed=3)
    full = _columns(ragged_batch_request(w))
    head = ragged_batch_request(w, count=10)
    sample = autotune._slice(10, full["group_offsets"], full["initial_spots"], full["strikes"],
                             full["maturities"], full["rates"], full["volatilities"], full["dividends"],
                             full["mesh"], full["path_counts"])
    for name in ("initial_spots", "strikes", "maturities", "rates", "volatilities"):
        assert np.array_equal(sample[name], head[name])
    assert np.array_equal(sample["group_offsets"], np.arange(11, dtype=np.uint32))
    for a, b in ((sample["dividends"].schedules, head["dividends"].schedules),
                 (sample["mesh"].schedules, head["mesh"].schedules)):
        assert np.array_equal(a.offsets - a.offsets[0], b.offsets - b.offsets[0])
        assert np.array_equal(a.times, b.times)
    assert np.array_equal(sample["dividends"].schedules.amounts, head["dividends"].schedules.amounts)
    assert np.array_equal(sample["path_counts"].training, head["path_counts"].training)
    # the method (stream seeds and domains) is shared unchanged: group g keeps domain + g
    assert head["method"].evaluation_stream.domain == 2 + 50


def test_sample_size_follows_the_path_work():
    from finoptions_hip import autotune
    from finoptions_hip.workloads import ragged_batch, ragged_batch_request

    w = ragged_batch(400, seed=5)
    c = _columns(ragged_batch_request(w))
    n = autotune._sample_groups(c["group_offsets"], c["mesh"], c["dividends"], None, c["path_counts"])
    rows = np.diff(c["mesh"].schedules.offsets) + 1
    work = rows * (c["path_counts"].training.astype(float) + c["path_counts"].evaluation) * 2.0
    assert 1 <= n < 400
    assert work[:n].sum() >= work.sum() / autotune.SAMPLE_FRACTION
    assert work[:n - 1].sum() < work.sum() / autotune.SAMPLE_FRACTION
