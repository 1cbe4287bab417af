#!/usr/bin/env python3
"""Headline benchmark: ragged-batch LSM path-steps/s including delta/vega, on one MI355X.

Workload (one step = one complete pricing request): the ragged batch of 10,000 Bermudan calls,
numpy default_rng(2024): exercise counts U{4..252} on an equal grid over T, 0-4 cash dividends at
exercise dates with amounts U[0.1, 2.0], S0 lognormal(ln 100, 0.3), K/S0 U[0.7, 1.3], sigma
U[0.1, 0.6], T U[0.1, 3.0], r U[0, 0.05], paths per contract log-uniform [16 384, 262 144]
(training = evaluation), Arrow-style int64 offsets, streams (2024, 1 + row) / (2024, 2 + rows +
row); the fused method (one HBM pass per exercise node, binary64 storage and decisions) with
pathwise delta + vega under the frozen policy, all contracts in one request.
Path-steps = intervals x (training + evaluation paths) per contract, the reference's unit.

Also reported: the plain FP64 method's price and price + bump-and-reprice delta/vega time on the
same batch (1 + 4 pricings), every fast price and Greek against the FP64 method in standard errors,
the binary32-storage method's time and accuracy, the fit step's roofline (HIP events on the
executor's stream against the algorithmic bytes 2 s N_train per node), the grouped-strike state
(64 strikes x 50 intervals, 2^20 + 2^20 paths) and the 4,096 x 256 grouped frame as extra keys,
and the CPU checker, contract-parallel like the reference, on a bounded sample of the batch.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "finoptions-engine_amd"))

METRIC = "LSM path-steps/sec incl. delta/vega vs plain FP64, 1 GPU, |bias|<0.1 SE"
HBM_PEAK_GBS = 8000.0
C4_CONTRACTS = 10_000
C2_STRIKES, C2_INTERVALS, C2_PATHS = 64, 50, 1 << 20
C5_GROUPS, C5_STRIKES, C5_PATHS = 4096, 256, 32_768
BUMP = 1e-4
# HBM bytes per fit-step launch of the headline pricing, measured with rocprofv3 FETCH_SIZE and
# WRITE_SIZE passes scaled by a copy-kernel calibration (2.000 x FETCH_SIZE, 1.000 x WRITE_SIZE
# KiB) and averaged over the pricing's 3011 launches: 902.3 MB read + 36.6 MB written.
MEASURED_TRAFFIC: dict[str, dict] = {
    "k_fit_bulk_step": {"bytes_per_launch": 938_900_000,
                        "source": "profiles/c4_ragged_batch_hbm_traffic.md"},
}


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


# ------------------------------------------------------------------------------ workloads

def c4_workload():
    from finoptions_hip.workloads import ragged_batch, ragged_path_steps

    w = ragged_batch(C4_CONTRACTS, seed=2024)
    return w, ragged_path_steps(w)


def price_c4(fh, executor, w, greeks=(), **bump):
    from finoptions_hip.workloads import ragged_batch_request

    return fh.price_lsm_call_batch(executor=executor, greeks=greeks,
                                   **ragged_batch_request(w, **bump))


def price_c2(fh, executor, greeks=(), *, spot=100.0, vol=0.25, paths=C2_PATHS):
    groups = fh.LsmStrikeGroups(
        group_offsets=np.array([0, C2_STRIKES], dtype=np.uint32),
        initial_spots=np.array([spot]), strikes=np.linspace(80.0, 120.0, C2_STRIKES),
        maturities=np.array([1.0]), rates=np.array([0.03]), volatilities=np.array([vol]),
        dividends=fh.ContinuousDividendYieldBatch(annualized_decimals=np.array([0.02])))
    mesh = fh.UniformExerciseMesh(base_intervals=C2_INTERVALS,
                                  origin_exercise=fh.OriginExercise.EXCLUDED)
    method = fh.Lsm(training_paths=paths, evaluation_paths=paths,
                    training_stream=fh.RandomStream(seed=11, domain=1),
                    evaluation_stream=fh.RandomStream(seed=11, domain=2),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    return fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=executor,
                                           method=method, greeks=greeks)


def c5_pricer(fh):
    from finoptions_hip.workloads import large_grouped_frame

    w = large_grouped_frame(groups=C5_GROUPS, strikes_per_group=C5_STRIKES)
    groups = fh.LsmStrikeGroups(
        group_offsets=w["group_offsets"], initial_spots=w["initial_spots"], strikes=w["strikes"],
        maturities=w["maturities"], rates=w["rates"], volatilities=w["volatilities"],
        dividends=fh.CashDividendBatch(
            schedules=fh.RaggedCashDividends(offsets=w["cash_offsets"], times=w["cash_times"],
                                             amounts=w["cash_amounts"]),
            policy=fh.DividendPolicy.LIQUIDATOR,
            maturity_boundary=fh.MaturityBoundary.STRICTLY_BEFORE))
    mesh = fh.UniformExerciseMesh(base_intervals=w["intervals"],
                                  origin_exercise=fh.OriginExercise.EXCLUDED)
    method = fh.Lsm(training_paths=C5_PATHS, evaluation_paths=C5_PATHS,
                    training_stream=fh.RandomStream(seed=5, domain=1),
                    evaluation_stream=fh.RandomStream(seed=5, domain=1 + C5_GROUPS),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.CONTINUE)
    steps = float(C5_GROUPS) * w["intervals"] * 2 * C5_PATHS

    def run(executor, greeks):
        return fh.price_lsm_call_grouped_batch(groups=groups, mesh=mesh, executor=executor,
                                               method=method, greeks=greeks)
    return run, steps


def timed(label, fn, steps, warmup):
    for i in range(warmup):
        fn()
        log(f"{label} warmup {i + 1}/{warmup} done")
    times, res = [], None
    for i in range(steps):
        t0 = time.perf_counter()
        res = fn()   # synchronous: returns after the device drains and the results are copied
        times.append(time.perf_counter() - t0)
        log(f"{label} step {i + 1}/{steps}: {times[-1] * 1e3:.1f} ms")
    return 1e3 * float(np.mean(times)), res


def bias_in_se(value, reference, se):
    se = np.asarray(se, dtype=np.float64)
    diff = np.abs(np.asarray(value) - np.asarray(reference))
    # a row whose every payoff is zero has SE 0 and must then agree exactly
    return np.where(se > 0, diff / np.where(se > 0, se, 1.0), np.where(diff == 0, 0.0, np.inf))


# ------------------------------------------------------------------------------ roofline

def fit_step_roofline(stats: dict, w: dict, real: int = 8) -> dict:
    """The fit step against HBM: algorithmic bytes 2 s N_train per active group and node (the
    node's share of the training matrix stored once and read once, 2 s M_obs N_train per group);
    achieved = those bytes over all launches / the launches' total time."""
    name = "k_fit_bulk_step"
    launches, total_ms = stats[name]
    fit_rows = np.diff(w["exercise_offsets"]).astype(np.float64)   # positive-time exercise rows
    alg_total = float(np.sum(2.0 * real * w["paths"].astype(np.float64) * fit_rows))
    achieved = alg_total / (total_ms / 1e3) / 1e9
    traffic = MEASURED_TRAFFIC.get(name)
    entry = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1),
             "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
             "traffic": traffic["bytes_per_launch"] if traffic else None,
             "launches": launches, "avg_launch_us": round(1e3 * total_ms / launches, 2),
             "algorithmic_bytes_per_launch": round(alg_total / launches),
             "algorithmic_bytes_definition": "2 x 8 B x N_train per active contract and node"}
    if traffic:
        entry["hbm_utilisation_gbs"] = round(
            traffic["bytes_per_launch"] / (1e-3 * total_ms / launches) / 1e9, 1)
        entry["traffic_source"] = traffic["source"]
    return entry


# ------------------------------------------------------------------------------ CPU baseline

def _oracle_block(pyoracle, w, count: int):
    """Contracts [0, count) of the ragged batch as one oracle row batch; group g keeps the domains
    of row g in the whole batch (training 1 + g, evaluation 2 + rows + g)."""
    rows = len(w["initial_spots"])
    eo, co = w["exercise_offsets"][:count + 1], w["cash_offsets"][:count + 1]
    return pyoracle.GroupedRequest(
        group_offsets=np.arange(count + 1, dtype=np.uint32),
        initial_spots=w["initial_spots"][:count], strikes=w["strikes"][:count],
        maturities=w["maturities"][:count], rates=w["rates"][:count],
        volatilities=w["volatilities"][:count], cash_offsets=co,
        cash_times=w["cash_times"][:co[-1]], cash_amounts=w["cash_amounts"][:co[-1]],
        cash_policy=0, training_paths=int(w["paths"][:count].max()),
        evaluation_paths=int(w["paths"][:count].max()),
        training_path_counts=w["paths"][:count], evaluation_path_counts=w["paths"][:count],
        training_seed=2024, evaluation_seed=2024, training_domain_start=1,
        evaluation_domain_start=2 + rows, mesh_kind=2, exercise_offsets=eo,
        exercise_times=w["exercise_times"][:eo[-1]], origin_exercise=0, insufficient_support=1)


def cpu_baseline(w, steps_per_contract, fast_res, plain_res, threads: int) -> dict:
    """The oracle, contract-parallel like the reference's row batch (pricing.rs:220-254: every
    contract serial, contracts spread over `threads` threads), timed on a bounded sample of the
    ragged batch (its first contracts, which are i.i.d. draws of the same generator), and on one
    thread over a smaller block; and, as the checker, the fast and plain GPU prices of the sampled
    contracts against the oracle's."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    cum = np.cumsum(steps_per_contract)
    count = int(np.searchsorted(cum, 1.0e9 * threads / 16)) + 1
    count_serial = int(np.searchsorted(cum, 1.0e8)) + 1
    t0 = time.perf_counter()
    out = pyoracle.price_grouped(_oracle_block(pyoracle, w, count), threads=threads)
    par_s = time.perf_counter() - t0
    t0 = time.perf_counter()
    pyoracle.price_grouped(_oracle_block(pyoracle, w, count_serial), threads=1)
    ser_s = time.perf_counter() - t0
    # the sanity case of docs/adr/0006-independent-policy-lsm.md:58-59 (M4 Pro: 101.31 ms at one
    # thread, 30.11 ms at four): 32 equal steps, logD = -0.06 t, cash 7.0 at node 16 (liquidator),
    # exercise node 16, sigma 0.30, S0 = K = 100, 65 536 + 65 536 paths, streams (10000, 1) /
    # (20000, 2) (crates/finoptions-lsm/benches/lsm_call.rs:22-120); the oracle's single call
    # simulates path-parallel and fits serially, as price_seeded does
    times = [k / 32 for k in range(33)]
    plan = pyoracle.Plan(times=times, log_discounts=[0.0] + [-0.06 * k / 32 for k in range(1, 33)],
                         cum_yields=[0.0] * 33, volatility=0.30, dividend_nodes=[16],
                         dividend_amounts=[7.0], exercise_nodes=[16], expiry_node=32)
    sanity = {}
    for th in (1, 4):
        pyoracle.set_threads(th)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            pyoracle.price_lsm_call(plan, 100.0, 100.0, 65_536, 65_536, 10_000, 1, 20_000, 2)
            best = min(best, time.perf_counter() - t0)
        sanity[f"ms_{th}_thread{'s' if th > 1 else ''}"] = round(1e3 * best, 2)
    sanity.update({"published_m4_pro_ms_1_thread": 101.31, "published_m4_pro_ms_4_threads": 30.11,
                   "workload": "32 steps, 65536 + 65536 paths, one exercise date, degree 4"})
    ref_p, ref_se = out["prices"], out["standard_errors"]
    fast_bias = bias_in_se(fast_res.prices[:count], ref_p, ref_se)
    plain_rel = float(np.max(np.abs(plain_res.prices[:count] - ref_p)
                             / np.maximum(np.abs(ref_p), 1e-300)))
    steps, steps_s = float(cum[count - 1]), float(cum[count_serial - 1])
    return {
        "value": steps / par_s, "unit": "path-steps/s", "cores": threads, "kind": "port",
        "sample": (f"contracts 0..{count - 1} of the ragged batch ({steps:.3e} path-steps), priced "
                   f"by the oracle contract-parallel over {threads} threads (each contract serial, "
                   f"as the reference's row batch), FP64 price only; {par_s:.1f} s"),
        "one_thread": {"value": steps_s / ser_s, "cores": 1,
                       "sample": f"contracts 0..{count_serial - 1} ({steps_s:.3e} path-steps), "
                                 f"{ser_s:.1f} s"},
        "sanity_32_step_case": sanity,
        "cpu": _cpu_model(),
        "checker_on_sample": {
            "contracts": count,
            "fast_max_price_bias_over_se": float(np.max(fast_bias)),
            "plain_fp64_max_price_rel_err": plain_rel,
            "plain_fp64_equal_stopping_counts": bool(
                np.array_equal(plain_res.expiry_paths[:count], out["expiry_paths"])),
        },
    }


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _c2_fixture() -> dict | None:
    path = ROOT / "tests" / "golden" / "c2_full_oracle.json"
    if not path.exists():
        return None
    raw = json.loads(path.read_text())
    return {k: (np.array([float.fromhex(x) for x in v]) if isinstance(v, list) and isinstance(v[0], str)
                else (np.array(v) if isinstance(v, list) else v)) for k, v in raw.items()}


def c1_leg(fh, fast, plain, threads: int) -> dict:
    """The single scheduled call (S0 = K = 100, times 0 / 0.5 / 1, cash 2.0 at node 1, exercise
    node 1, 65 536 + 65 536 paths): the fused call with all four Greeks, the plain FP64 call with
    4 bump pricings of delta and vega, and the oracle's call on 1 and `threads` threads."""
    schedule = fh.EventSchedule(times=(0.0, 0.5, 1.0), dividends=(fh.CashDividend(node=1, amount=2.0),),
                                exercise_nodes=(1,), expiry_node=2)
    curve = fh.LogDiscountCurve(times=schedule.times, log_discounts=(0.0, -0.01, -0.025))
    method = fh.Lsm(training_paths=65_536, evaluation_paths=65_536,
                    training_stream=fh.RandomStream(seed=7, domain=1),
                    evaluation_stream=fh.RandomStream(seed=7, domain=2),
                    insufficient_exercise_support=fh.InsufficientExerciseSupportPolicy.FAIL)

    def call(executor, spot=100.0, vol=0.20, greeks=()):
        return fh.price_lsm_call(initial_spot=spot, strike=100.0, schedule=schedule, curve=curve,
                                 dividend_yield_curve=fh.CumulativeDividendYieldCurve.zero(times=schedule.times),
                                 volatility=fh.ConstantVolatility(vol), dividend_policy=fh.DividendPolicy.LIQUIDATOR,
                                 executor=executor, method=method, greeks=greeks)

    all4 = (fh.Greek.DELTA, fh.Greek.VEGA, fh.Greek.RHO, fh.Greek.DIVIDEND_YIELD)
    for _ in range(5):
        call(fast, greeks=all4)
    reps = 50
    t0 = time.perf_counter()
    for _ in range(reps):
        res = call(fast, greeks=all4)
    ms_fast = 1e3 * (time.perf_counter() - t0) / reps
    out = {"workload": "one call: S0 = K = 100, times (0, 0.5, 1), cash 2.0 at node 1 (liquidator), "
                       "exercise node 1, logD (0, -0.01, -0.025), sigma 0.2, 65536 + 65536 paths, "
                       "streams (7, 1) / (7, 2); 262144 path-steps",
           "fast_ms_with_4_greeks": round(ms_fast, 3), "price": res.price}
    if plain is not None:
        for _ in range(3):
            call(plain)
        t0 = time.perf_counter()
        for _ in range(10):
            call(plain)
            for kw in ({"spot": 100 * (1 + BUMP)}, {"spot": 100 * (1 - BUMP)},
                       {"vol": 0.20 * (1 + BUMP)}, {"vol": 0.20 * (1 - BUMP)}):
                call(plain, **kw)
        ms_plain = 1e3 * (time.perf_counter() - t0) / 10
        out["plain_ms_price_bump_greeks"] = round(ms_plain, 3)
        out["fast_speedup_vs_plain_price_bump_greeks"] = round(ms_plain / ms_fast, 2)
    sys.path.insert(0, str(ROOT / "oracle"))
    import pyoracle

    plan = pyoracle.Plan(times=[0.0, 0.5, 1.0], log_discounts=[0.0, -0.01, -0.025], cum_yields=[0.0, 0.0, 0.0],
                         volatility=0.20, dividend_policy=0, dividend_nodes=[1], dividend_amounts=[2.0],
                         exercise_nodes=[1], expiry_node=2)
    for th in sorted({1, threads}):
        pyoracle.set_threads(th)
        best = float("inf")
        for _ in range(3):
            t0 = time.perf_counter()
            ref = pyoracle.price_lsm_call(plan, 100.0, 100.0, 65_536, 65_536, 7, 1, 7, 2)
            best = min(best, time.perf_counter() - t0)
        out[f"cpu_oracle_ms_{th}_thread{'s' if th > 1 else ''}"] = round(1e3 * best, 2)
    out["oracle_price"] = ref["price"]
    out["fast_price_rel_err_vs_oracle"] = abs(res.price - ref["price"]) / ref["price"]
    return out


def copy_calibration() -> int:
    """Known-size traffic for the HBM counters: each launch reads and writes exactly 256 MiB."""
    import torch

    n = 64 << 20
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        torch.add(a, 0.0, out=b)
    torch.cuda.synchronize()
    log(f"copy calibration: 3 x {4 * n >> 20} MiB read and written; check {float(b[0])}")
    return 0


# ------------------------------------------------------------------------------ main

def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--skip-cpu", action="store_true")
    ap.add_argument("--skip-plain", action="store_true")
    ap.add_argument("--skip-extras", action="store_true",
                    help="skip the grouped-strike state, the grouped frame and the FP32-bulk method")
    ap.add_argument("--only", choices=("c4", "c2", "c5"),
                    help="time one workload's fused pricing (with delta/vega) and print a short "
                         "line: c4 the ragged batch, c2 the grouped-strike state, c5 the grouped "
                         "frame (for profiler runs)")
    ap.add_argument("--fit-cashflows", choices=("auto", "value_columns", "stopping_rows"),
                    default="auto", help="cashflow encoding of the fused fit step")
    ap.add_argument("--batch-schedule", choices=("serial", "overlapped"), default="serial",
                    help="how the memory-bounded batches of a request share the device")
    ap.add_argument("--fit-schedule", choices=("auto", "per_node", "sweep"), default="auto",
                    help="launch form of the fused backward induction: one fit pass per exercise "
                         "row, or one persistent sweep per batch (single-strike groups)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the reference copy for the rocprofv3 FETCH_SIZE / WRITE_SIZE scale "
                         "factors: three elementwise copies of 256 MiB (64 Mi float32)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        log("bench.py measures one GPU only: every box in this setup has a single MI355X and the "
            "multi-GPU sharding of this workload is a design note, not a measured path. "
            "Run with --gpus 1.")
        return 2

    import finoptions_hip as fh

    fast = fh.HipExecutor(device=0, precision=fh.HipPrecision.FUSED_F64)
    fast.tune(fit_cashflows=args.fit_cashflows, batch_schedule=args.batch_schedule,
              fit_schedule=args.fit_schedule)
    greeks = (fh.Greek.DELTA, fh.Greek.VEGA)
    if args.only in ("c2", "c5"):
        if args.only == "c2":
            fn, steps = (lambda: price_c2(fh, fast, greeks)), C2_INTERVALS * 2.0 * C2_PATHS
        else:
            run5, steps = c5_pricer(fh)
            fn = lambda: run5(fast, greeks)   # noqa: E731
        ms, _ = timed(args.only + " fast", fn, args.steps, args.warmup)
        print(json.dumps({"workload": args.only, "fit_cashflows": args.fit_cashflows,
                          "path_steps_per_s": steps / (ms / 1e3), "ms_per_step": ms}), flush=True)
        return 0
    w, steps_per = c4_workload()
    path_steps = float(np.sum(steps_per))
    log(f"device: {fast.device_name}; ragged batch of {C4_CONTRACTS} contracts, "
        f"{path_steps:.4e} path-steps per pricing")

    # ---- headline: fused method + pathwise delta/vega on the whole ragged batch
    ms_fast, fast_res = timed("c4 fast", lambda: price_c4(fh, fast, w, greeks), args.steps,
                              args.warmup)
    value = path_steps / (ms_fast / 1e3)
    if args.only == "c4":
        print(json.dumps({"workload": "c4", "fit_cashflows": args.fit_cashflows,
                          "path_steps_per_s": value, "ms_per_step": ms_fast}), flush=True)
        return 0

    # ---- per-kernel times of one headline pricing, HIP events on the executor's stream
    fast.profile(True)
    price_c4(fh, fast, w, greeks)
    stats = fast.kernel_stats()
    fast.profile(False)
    kernel_total = sum(v[1] for v in stats.values())
    kernel_table = {k: {"launches": n, "total_ms": round(ms, 3), "avg_us": round(1e3 * ms / n, 2),
                        "share": round(ms / kernel_total, 3)}
                    for k, (n, ms) in sorted(stats.items(), key=lambda kv: -kv[1][1])}
    log("kernel times (one headline pricing): " + json.dumps(kernel_table))
    dominant = max(stats.items(), key=lambda kv: kv[1][1])[0]
    roofline = fit_step_roofline(stats, w)
    roofline["dominant_kernel"] = dominant
    roofline["dominant_share"] = kernel_table[dominant]["share"]
    generation = {}
    for k in ("k_simulate", "k_evaluate"):
        if k in stats:
            # one simulated path-step per path and interval of the population the kernel covers
            generation[k] = {"path_steps_per_s": round((path_steps / 2) / (stats[k][1] / 1e3)),
                             "total_ms": round(stats[k][1], 2)}
    roofline["generation_kernels"] = generation

    # ---- plain FP64: price + bump-and-reprice delta/vega (1 + 4 pricings) on the same batch
    plain_info = accuracy = None
    plain_res = None
    if not args.skip_plain:
        plain = fh.HipExecutor(device=0, precision=fh.HipPrecision.F64)
        price_c4(fh, plain, w)   # allocation warmup for this executor
        t0 = time.perf_counter()
        plain_res = price_c4(fh, plain, w)
        t_price = time.perf_counter() - t0
        t0 = time.perf_counter()
        up_s = price_c4(fh, plain, w, spot_scale=1 + BUMP)
        dn_s = price_c4(fh, plain, w, spot_scale=1 - BUMP)
        up_v = price_c4(fh, plain, w, vol_shift=BUMP)
        dn_v = price_c4(fh, plain, w, vol_shift=-BUMP)
        t_bumps = time.perf_counter() - t0
        ms_plain = 1e3 * (t_price + t_bumps)
        log(f"plain FP64: price {t_price:.2f} s, price + 4 bumps {ms_plain / 1e3:.2f} s")
        ref = price_c4(fh, plain, w, greeks)   # FP64 pathwise Greeks: the accuracy reference
        bp = bias_in_se(fast_res.prices, plain_res.prices, plain_res.standard_errors)
        bd = bias_in_se(fast_res.greeks[fh.Greek.DELTA], ref.greeks[fh.Greek.DELTA],
                        ref.greek_standard_errors[fh.Greek.DELTA])
        bv = bias_in_se(fast_res.greeks[fh.Greek.VEGA], ref.greeks[fh.Greek.VEGA],
                        ref.greek_standard_errors[fh.Greek.VEGA])
        accuracy = {
            "reference": "plain FP64 GPU method on all 10000 contracts (it matches the oracle to "
                         "1e-10 relative, see cpu_baseline.checker_on_sample)",
            "max_price_bias_over_se": float(bp.max()),
            "max_delta_bias_over_se": float(bd.max()),
            "max_vega_bias_over_se": float(bv.max()),
            "all_within_0.1_se": bool(max(bp.max(), bd.max(), bv.max()) < 0.1),
        }
        plain_info = {"ms_price": round(t_price * 1e3, 1),
                      "ms_price_bump_greeks": round(ms_plain, 1),
                      "fast_speedup_vs_plain_price_bump_greeks": round(ms_plain / ms_fast, 2),
                      "plain_path_steps_per_s_price_bump_greeks": path_steps * 5 / (ms_plain / 1e3)}
        del up_s, dn_s, up_v, dn_v

    extras = {}
    if not args.skip_extras:
        # ---- binary32 spots and cashflows (FP64 Gram/solve) on the same batch
        bulk = fh.HipExecutor(device=0, precision=fh.HipPrecision.F32_BULK)
        ms_b, bulk_res = timed("c4 f32-bulk", lambda: price_c4(fh, bulk, w, greeks), 2, 1)
        f32 = {"ms_per_step": round(ms_b, 1), "path_steps_per_s": path_steps / (ms_b / 1e3)}
        if plain_res is not None:
            b = max(float(bias_in_se(bulk_res.prices, plain_res.prices,
                                     plain_res.standard_errors).max()),
                    float(bias_in_se(bulk_res.greeks[fh.Greek.DELTA], ref.greeks[fh.Greek.DELTA],
                                     ref.greek_standard_errors[fh.Greek.DELTA]).max()))
            f32["max_bias_over_se"] = b
            f32["meets_0.1_se_bar"] = bool(b < 0.1)
        extras["c4_f32_bulk"] = f32
        del bulk
        # ---- grouped-strike state at full size
        c2_steps = C2_INTERVALS * 2 * C2_PATHS
        ms_c2, c2_res = timed("c2 fast", lambda: price_c2(fh, fast, greeks), 10, 2)
        c2 = {"workload": "1 state x 64 strikes linspace(80,120), S0 100, r 0.03, q 0.02, sigma "
                          "0.25, T 1, 50 intervals, 2^20 + 2^20 paths, seed 11 domains 1/2",
              "ms_per_step": round(ms_c2, 3), "path_steps_per_s": c2_steps / (ms_c2 / 1e3),
              "contract_equivalent_path_steps_per_s": C2_STRIKES * c2_steps / (ms_c2 / 1e3)}
        if not args.skip_plain:
            price_c2(fh, plain)
            t0 = time.perf_counter()
            base2 = price_c2(fh, plain)
            for kw in ({"spot": 100 * (1 + BUMP)}, {"spot": 100 * (1 - BUMP)},
                       {"vol": 0.25 * (1 + BUMP)}, {"vol": 0.25 * (1 - BUMP)}):
                price_c2(fh, plain, **kw)
            ms_p2 = 1e3 * (time.perf_counter() - t0)
            c2["plain_ms_price_bump_greeks"] = round(ms_p2, 1)
            c2["fast_speedup_vs_plain_price_bump_greeks"] = round(ms_p2 / ms_c2, 2)
            fx = _c2_fixture()
            if fx is not None:
                c2["fast_max_price_bias_over_se_vs_oracle"] = float(
                    bias_in_se(c2_res.prices, fx["prices"], fx["standard_errors"]).max())
                c2["plain_max_price_rel_err_vs_oracle"] = float(
                    np.max(np.abs(base2.prices - fx["prices"]) / fx["prices"]))
                c2["oracle_reference"] = "tests/golden/c2_full_oracle.json"
        extras["c2_grouped_strikes"] = c2
        # ---- the 4096 x 256 grouped frame
        run5, steps5 = c5_pricer(fh)
        ms_c5, _ = timed("c5 fast", lambda: run5(fast, greeks), 2, 1)
        extras["c5_grouped_frame"] = {
            "workload": "4096 states x 256 strikes, 12 monthly nodes, quarterly cash, 32768 + "
                        "32768 paths per state",
            "ms_per_step": round(ms_c5, 1), "path_steps_per_s": steps5 / (ms_c5 / 1e3),
            "strike_rows_per_s": C5_GROUPS * C5_STRIKES / (ms_c5 / 1e3)}
        # ---- the single scheduled call
        c1_threads = max(1, min(16, len(os.sched_getaffinity(0))))
        extras["c1_single_call"] = c1_leg(fh, fast, None if args.skip_plain else plain, c1_threads)

    cpu = None
    if not args.skip_cpu and plain_res is not None:
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        log(f"timing the CPU checker on a bounded sample over {threads} threads ...")
        cpu = cpu_baseline(w, steps_per, fast_res, plain_res, threads)
        cpu["fast_speedup_vs_cpu"] = round(value / cpu["value"], 1)

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "path-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_fast,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": (f"ragged batch of {C4_CONTRACTS} Bermudan calls, seed 2024: exercise "
                         "counts U{4..252}, 0-4 cash dividends U[0.1,2.0] at exercise dates, S0 "
                         "lognormal(ln100,0.3), K/S0 U[0.7,1.3], sigma U[0.1,0.6], T U[0.1,3], r "
                         "U[0,0.05], paths log-uniform [16384,262144] (train = eval), Arrow "
                         "offsets, one request"),
            "method": ("fused: one HBM pass per exercise node (binary64 spots and cashflows, "
                       "pivoted power sums, moment-form Stieltjes solve), binary64 exercise "
                       "decisions, binary32 Greek tangents"),
            "greeks": "pathwise delta+vega, frozen policy, evaluation paths regenerated from Philox",
            "path_steps_per_step": path_steps,
            "parallelism": "single GPU",
        },
        "roofline": roofline,
        "cpu_baseline": cpu,
        "plain_fp64": plain_info,
        "accuracy": accuracy,
        "kernels": kernel_table,
        **extras,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
