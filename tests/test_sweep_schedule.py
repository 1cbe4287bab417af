"""The fit sweep's schedule (host logic, no device): every group with a backward step is swept by
one team whose members hold disjoint runs of at most `chunks_per_block` chunks covering all of
the group's chunks; teams come in order of decreasing cost with each team's members on
consecutive tickets; and a discrete-event run, where R resident blocks take tickets one after
another, a team steps only once all its members have joined and durations are random, always
finishes when R is at least the largest team (and stalls when it is not: the run is a real
check)."""

from __future__ import annotations

import ctypes as C
import heapq

import numpy as np
import pytest


def _schedule(chunks, fit_rows, q):
    from finoptions_hip._lib import lib

    chunks = np.ascontiguousarray(chunks, dtype=np.uint32)
    fit_rows = np.ascontiguousarray(fit_rows, dtype=np.uint32)
    u32p = C.POINTER(C.c_uint32)
    count, teams = C.c_uint64(0), C.c_uint32(0)
    code = lib().fohip_sweep_schedule(chunks.ctypes.data_as(u32p), fit_rows.ctypes.data_as(u32p), len(chunks),
                                      q, None, 0, C.byref(count), C.byref(teams))
    if count.value == 0:
        return code, np.zeros((0, 8), dtype=np.uint32), teams.value
    entries = np.zeros((count.value, 8), dtype=np.uint32)
    code = lib().fohip_sweep_schedule(chunks.ctypes.data_as(u32p), fit_rows.ctypes.data_as(u32p), len(chunks),
                                      q, entries.ctypes.data_as(u32p), count.value, C.byref(count),
                                      C.byref(teams))
    return code, entries, teams.value


def _check(chunks, fit_rows, q):
    code, entries, teams = _schedule(chunks, fit_rows, q)
    assert code == 0
    group, c0, c1, members, team, leader = (entries[:, i].astype(np.int64) for i in range(6))
    part = entries[:, 6].astype(np.uint64) | (entries[:, 7].astype(np.uint64) << np.uint64(32))
    swept = [g for g in range(len(chunks)) if fit_rows[g] > 0 and chunks[g] > 0]
    assert sorted(set(group.tolist())) == swept and teams == len(swept)
    # ticket order: teams 0, 1, ... each on consecutive tickets, members in chunk order, leader first
    assert team[0] == 0 and np.all(np.diff(team) >= 0) and np.all(np.diff(team) <= 1)
    spans, costs = [], []
    for t in range(teams):
        mine = np.flatnonzero(team == t)
        g = int(group[mine[0]])
        m = -(-int(chunks[g]) // q)
        assert np.all(group[mine] == g) and len(mine) == m and np.all(members[mine] == m)
        assert c0[mine[0]] == 0 and c1[mine[-1]] == chunks[g] and np.all(c0[mine[1:]] == c1[mine[:-1]])
        assert np.all((c1[mine] - c0[mine] > 0) & (c1[mine] - c0[mine] <= q))
        assert leader[mine].tolist() == [1] + [0] * (m - 1)
        assert len(set(part[mine].tolist())) == 1
        spans.append((int(part[mine[0]]), int(part[mine[0]]) + 2 * int(chunks[g])))
        costs.append(int(fit_rows[g]) * (-(-int(chunks[g]) // m) + 4.0))
    spans.sort()
    assert all(a[1] <= b[0] for a, b in zip(spans, spans[1:]))   # disjoint partial records
    assert all(a >= b for a, b in zip(costs, costs[1:]))          # longest sweeps first
    return entries, teams


def _run(entries, resident, rng):
    """`resident` blocks take tickets in order whenever they are free; a team runs once all its
    members have joined, for a random time per member. True when every ticket is done."""
    members = entries[:, 3].astype(int)
    team = entries[:, 4].astype(int)
    nxt, joined, done = 0, {}, 0
    free = [(0.0, b) for b in range(resident)]
    heapq.heapify(free)
    while free:
        now, b = heapq.heappop(free)
        if nxt >= len(entries):
            continue
        e = nxt
        nxt += 1
        t = team[e]
        joined.setdefault(t, []).append(b)
        if len(joined[t]) == members[e]:
            length = float(rng.exponential(1.0) * (1 + entries[e, 2] - entries[e, 1]))
            for m in joined.pop(t):
                done += 1
                heapq.heappush(free, (now + length * rng.uniform(0.5, 1.5), m))
    return done == len(entries) and not joined


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_schedule_covers_every_group_once_and_always_completes(seed):
    rng = np.random.default_rng(seed)
    groups = 400
    chunks = rng.integers(1, 140, groups)
    fit_rows = rng.integers(0, 253, groups)
    fit_rows[rng.random(groups) < 0.05] = 0
    for q in (3, 5, 1):
        entries, teams = _check(chunks, fit_rows, q)
        most = int(entries[:, 3].max())
        for resident in (most, most + 7, 512):
            assert _run(entries, resident, rng)


def test_ragged_batch_schedule_completes_on_few_blocks():
    """The headline ragged batch's contracts (8 to 128 chunks, 4 to 252 steps): teams of 3 to 43
    blocks; any residency from the largest team up completes."""
    from finoptions_hip.workloads import ragged_batch

    w = ragged_batch(2000)
    chunks = -(-w["paths"].astype(np.int64) // 2048)
    fit_rows = np.diff(w["exercise_offsets"])
    entries, teams = _check(chunks, fit_rows, 3)
    most = int(entries[:, 3].max())
    assert most <= 256   # within one block per CU of an MI355X
    rng = np.random.default_rng(0)
    for resident in (most, 256, 512):
        assert _run(entries, resident, rng)


def test_small_inputs_and_the_run_detects_a_stall():
    entries, teams = _check([1, 2, 3], [5, 1, 2], 3)
    assert teams == 3 and len(entries) == 3
    code, entries, teams = _schedule([4, 5], [0, 0], 3)
    assert code == 0 and len(entries) == 0 and teams == 0
    # fewer resident blocks than the largest team: the run stalls, as the device would
    entries, _ = _check([30, 2], [4, 4], 3)
    assert int(entries[:, 3].max()) == 10
    assert not _run(entries, 9, np.random.default_rng(0))
    assert _run(entries, 10, np.random.default_rng(0))
