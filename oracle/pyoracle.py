"""ctypes bindings of the CPU parity oracle (oracle/lsm_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()`` as the checker and by
``bench.py``'s ``cpu_baseline`` leg. The product package never imports this module.
"""

from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "lib" / "liblsm_oracle.so"


def build() -> Path:
    """Compile the oracle with its Makefile (gcc, -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


class OrError(C.Structure):
    _fields_ = [
        ("code", C.c_int32),
        ("phase", C.c_int32),
        ("row", C.c_int64),
        ("node", C.c_int64),
        ("path", C.c_int64),
        ("index", C.c_int64),
        ("observed", C.c_double),
        ("limit", C.c_double),
    ]


class OrPlanInput(C.Structure):
    _fields_ = [
        ("nodes", C.c_uint32),
        ("times", C.POINTER(C.c_double)),
        ("log_discounts", C.POINTER(C.c_double)),
        ("cum_yields", C.POINTER(C.c_double)),
        ("volatility", C.c_double),
        ("dividend_policy", C.c_int32),
        ("dividends", C.c_uint32),
        ("dividend_nodes", C.POINTER(C.c_uint32)),
        ("dividend_amounts", C.POINTER(C.c_double)),
        ("exercise_count", C.c_uint32),
        ("exercise_nodes", C.POINTER(C.c_uint32)),
        ("expiry_node", C.c_uint32),
    ]


class OrModel(C.Structure):
    _fields_ = [
        ("degree", C.c_int32),
        ("mean", C.c_double),
        ("scale", C.c_double),
        ("alpha", C.c_double * 4),
        ("beta", C.c_double * 4),
        ("coef", C.c_double * 5),
        ("norm", C.c_double * 5),
        ("observations", C.c_uint32),
        ("rank", C.c_uint32),
        ("condition", C.c_double),
    ]


class OrEstimator(C.Structure):
    _fields_ = [
        ("strike", C.c_double),
        ("degree", C.c_int32),
        ("rank_tolerance", C.c_double),
        ("condition_limit", C.c_double),
        ("tie_rule", C.c_int32),
        ("insufficient_support", C.c_int32),
        ("tie_window", C.c_double),
    ]


class OrNodeFit(C.Structure):
    _fields_ = [
        ("node", C.c_uint32),
        ("kind", C.c_int32),
        ("exercise", C.c_int32),
        ("estimated_continuation", C.c_double),
        ("itm_paths", C.c_uint32),
        ("required_paths", C.c_uint32),
        ("model", OrModel),
        ("near_ties", C.c_uint32),
    ]


class OrCallResult(C.Structure):
    _fields_ = [
        ("price", C.c_double),
        ("paths", C.c_uint32),
        ("has_sd", C.c_int32),
        ("sd", C.c_double),
        ("se", C.c_double),
        ("obs", C.c_uint32),
        ("observation_nodes", C.POINTER(C.c_uint32)),
        ("exercise_counts", C.POINTER(C.c_uint32)),
        ("fits", C.POINTER(OrNodeFit)),
        ("fit_count", C.c_uint32),
        ("near_tie_training", C.c_uint64),
        ("near_tie_evaluation", C.c_uint64),
    ]


class OrGroupedInput(C.Structure):
    _fields_ = [
        ("groups", C.c_uint32),
        ("group_offsets", C.POINTER(C.c_uint32)),
        ("initial_spots", C.POINTER(C.c_double)),
        ("maturities", C.POINTER(C.c_double)),
        ("rates", C.POINTER(C.c_double)),
        ("volatilities", C.POINTER(C.c_double)),
        ("rows", C.c_uint32),
        ("strikes", C.POINTER(C.c_double)),
        ("dividend_model", C.c_int32),
        ("dividend_yields", C.POINTER(C.c_double)),
        ("cash_offsets", C.POINTER(C.c_int64)),
        ("cash_times", C.POINTER(C.c_double)),
        ("cash_amounts", C.POINTER(C.c_double)),
        ("cash_policy", C.c_int32),
        ("mesh_kind", C.c_int32),
        ("base_intervals", C.c_uint32),
        ("origin_exercise", C.c_int32),
        ("exercise_offsets", C.POINTER(C.c_int64)),
        ("exercise_times", C.POINTER(C.c_double)),
        ("training_paths", C.c_uint32),
        ("evaluation_paths", C.c_uint32),
        ("training_path_counts", C.POINTER(C.c_uint32)),
        ("evaluation_path_counts", C.POINTER(C.c_uint32)),
        ("training_seed", C.c_uint64),
        ("evaluation_seed", C.c_uint64),
        ("training_domain_start", C.c_uint32),
        ("evaluation_domain_start", C.c_uint32),
        ("degree", C.c_int32),
        ("rank_tolerance", C.c_double),
        ("condition_limit", C.c_double),
        ("tie_rule", C.c_int32),
        ("insufficient_support", C.c_int32),
        ("escrowed_volatility", C.c_int32),
        ("tie_window", C.c_double),
    ]


class OrBatchOutput(C.Structure):
    _fields_ = [
        ("prices", C.POINTER(C.c_double)),
        ("sds", C.POINTER(C.c_double)),
        ("ses", C.POINTER(C.c_double)),
        ("origin_fitted", C.POINTER(C.c_uint8)),
        ("origin_exercised", C.POINTER(C.c_uint8)),
        ("regression_nodes", C.POINTER(C.c_uint32)),
        ("no_support_nodes", C.POINTER(C.c_uint32)),
        ("insufficient_nodes", C.POINTER(C.c_uint32)),
        ("min_itm_paths", C.POINTER(C.c_uint32)),
        ("max_condition", C.POINTER(C.c_double)),
        ("early_paths", C.POINTER(C.c_uint32)),
        ("expiry_paths", C.POINTER(C.c_uint32)),
        ("near_tie_training", C.POINTER(C.c_uint64)),
        ("near_tie_evaluation", C.POINTER(C.c_uint64)),
    ]


_lib: C.CDLL | None = None


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.or_price_grouped.argtypes = [
            C.POINTER(OrGroupedInput), C.POINTER(OrBatchOutput), C.c_int, C.POINTER(OrError)]
    return _lib


def _p(arr: np.ndarray, ctype):
    return arr.ctypes.data_as(C.POINTER(ctype))


class OracleError(RuntimeError):
    def __init__(self, err: OrError):
        self.code = int(err.code)
        self.phase = int(err.phase)
        self.row = int(err.row)
        self.node = int(err.node)
        self.path = int(err.path)
        self.index = int(err.index)
        self.observed = float(err.observed)
        self.limit = float(err.limit)
        super().__init__(
            f"oracle error {self.code} phase={self.phase} row={self.row} node={self.node} "
            f"path={self.path} index={self.index} observed={self.observed} limit={self.limit}")


# ------------------------------------------------------------------------------------ random


def philox(counter, key) -> np.ndarray:
    c = np.asarray(counter, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().or_philox4x32_10(_p(c, C.c_uint32), _p(k, C.c_uint32), _p(out, C.c_uint32))
    return out


def stream_u64(seed: int, domain: int, path: int, count: int) -> np.ndarray:
    out = np.zeros(count, dtype=np.uint64)
    lib().or_stream_u64(C.c_uint64(seed), C.c_uint32(domain), C.c_uint32(path),
                        C.c_uint64(count), _p(out, C.c_uint64))
    return out


def normals(seed: int, domain: int, path: int, count: int) -> tuple[np.ndarray, np.ndarray]:
    out = np.zeros(count, dtype=np.float64)
    draws = np.zeros(count, dtype=np.uint32)
    rc = lib().or_stream_normals(C.c_uint64(seed), C.c_uint32(domain), C.c_uint32(path),
                                 C.c_uint32(count), _p(out, C.c_double), _p(draws, C.c_uint32))
    if rc:
        raise RuntimeError(f"normal generation failed with {rc}")
    return out, draws


def uniforms(bits: np.ndarray) -> dict[str, np.ndarray]:
    b = np.ascontiguousarray(bits, dtype=np.uint64)
    res = {}
    for name in ("symmetric_unit", "open_unit", "standard_unit"):
        out = np.zeros(b.size, dtype=np.float64)
        getattr(lib(), "or_" + name)(_p(b, C.c_uint64), C.c_size_t(b.size), _p(out, C.c_double))
        res[name] = out
    return res


def zig_tables() -> tuple[np.ndarray, np.ndarray, float]:
    x = np.zeros(257)
    f = np.zeros(257)
    r = C.c_double()
    lib().or_zig_tables(_p(x, C.c_double), _p(f, C.c_double), C.byref(r))
    return x, f, r.value


# ------------------------------------------------------------------------------------ plans


@dataclass
class Plan:
    """Explicit event schedule + curves + volatility of one GBM path plan."""

    times: list[float]
    log_discounts: list[float]
    cum_yields: list[float]
    volatility: float
    dividend_policy: int = 0
    dividend_nodes: list[int] = field(default_factory=list)
    dividend_amounts: list[float] = field(default_factory=list)
    exercise_nodes: list[int] = field(default_factory=list)
    expiry_node: int | None = None

    def struct(self) -> tuple[OrPlanInput, list]:
        keep = [
            np.ascontiguousarray(self.times, dtype=np.float64),
            np.ascontiguousarray(self.log_discounts, dtype=np.float64),
            np.ascontiguousarray(self.cum_yields, dtype=np.float64),
            np.ascontiguousarray(self.dividend_nodes or [0], dtype=np.uint32),
            np.ascontiguousarray(self.dividend_amounts or [0.0], dtype=np.float64),
            np.ascontiguousarray(self.exercise_nodes or [0], dtype=np.uint32),
        ]
        expiry = len(self.times) - 1 if self.expiry_node is None else self.expiry_node
        s = OrPlanInput(
            len(self.times), _p(keep[0], C.c_double), _p(keep[1], C.c_double),
            _p(keep[2], C.c_double), float(self.volatility), int(self.dividend_policy),
            len(self.dividend_nodes), _p(keep[3], C.c_uint32), _p(keep[4], C.c_double),
            len(self.exercise_nodes), _p(keep[5], C.c_uint32), expiry)
        return s, keep

    @property
    def observations(self) -> int:
        return len(self.exercise_nodes) + 1


def step_coefficients(plan: Plan) -> tuple[np.ndarray, np.ndarray]:
    s, keep = plan.struct()
    n = len(plan.times) - 1
    drift = np.zeros(n)
    diff = np.zeros(n)
    lib().or_step_coefficients(C.byref(s), _p(drift, C.c_double), _p(diff, C.c_double))
    return drift, diff


def simulate_seeded(plan: Plan, initial_spot: float, paths: int, seed: int,
                    domain: int) -> np.ndarray:
    s, keep = plan.struct()
    out = np.zeros((plan.observations, paths), dtype=np.float64)
    err = OrError()
    rc = lib().or_simulate_seeded(C.byref(s), C.c_double(initial_spot), C.c_uint32(paths),
                                  C.c_uint64(seed), C.c_uint32(domain), _p(out, C.c_double),
                                  C.byref(err))
    if rc:
        raise OracleError(err)
    return out


def simulate_normals(plan: Plan, initial_spot: float, normals_matrix: np.ndarray) -> np.ndarray:
    s, keep = plan.struct()
    z = np.ascontiguousarray(normals_matrix, dtype=np.float64)
    paths = z.shape[1]
    out = np.zeros((plan.observations, paths), dtype=np.float64)
    err = OrError()
    rc = lib().or_simulate_normals(C.byref(s), C.c_double(initial_spot), C.c_uint32(paths),
                                   _p(z, C.c_double), _p(out, C.c_double), C.byref(err))
    if rc:
        raise OracleError(err)
    return out


# ------------------------------------------------------------------------------- regression


def model_dict(m: OrModel) -> dict:
    return {
        "degree": m.degree, "mean": m.mean, "scale": m.scale,
        "alpha": list(m.alpha), "beta": list(m.beta), "coef": list(m.coef),
        "norm": list(m.norm), "observations": m.observations, "rank": m.rank,
        "condition": m.condition,
    }


def fit_regression(x, y, degree: int = 4, rank_tolerance: float = 1e-12,
                   condition_limit: float = 1e8) -> tuple[dict, OrModel]:
    xa = np.ascontiguousarray(x, dtype=np.float64)
    ya = np.ascontiguousarray(y, dtype=np.float64)
    m = OrModel()
    err = OrError()
    rc = lib().or_fit_regression(_p(xa, C.c_double), _p(ya, C.c_double), C.c_uint32(xa.size),
                                 C.c_int32(degree), C.c_double(rank_tolerance),
                                 C.c_double(condition_limit), C.byref(m), C.byref(err))
    if rc:
        raise OracleError(err)
    return model_dict(m), m


def predict(model: OrModel, x) -> np.ndarray:
    f = lib().or_predict
    f.restype = C.c_double
    return np.array([f(C.byref(model), C.c_double(float(v))) for v in np.atleast_1d(x)])


def polynomial_moments(z, y, b) -> tuple[float, float, float]:
    za, ya, ba = (np.ascontiguousarray(v, dtype=np.float64) for v in (z, y, b))
    out = np.zeros(3)
    lib().or_polynomial_moments(_p(za, C.c_double), _p(ya, C.c_double), _p(ba, C.c_double),
                                C.c_size_t(za.size), _p(out, C.c_double))
    return tuple(out)


def recurrence_moments(z, y, prev, cur, alpha, beta):
    za, ya, pa, ca = (np.ascontiguousarray(v, dtype=np.float64) for v in (z, y, prev, cur))
    nxt = np.zeros(za.size)
    out = np.zeros(3)
    lib().or_recurrence_moments(_p(za, C.c_double), _p(ya, C.c_double), _p(pa, C.c_double),
                                _p(ca, C.c_double), _p(nxt, C.c_double), C.c_size_t(za.size),
                                C.c_double(alpha), C.c_double(beta), _p(out, C.c_double))
    return tuple(out), nxt


# ------------------------------------------------------------------------------------- LSM

KIND_NAMES = {0: "origin_decision", 1: "no_in_the_money_paths",
              2: "insufficient_exercise_support", 3: "regression"}


def _call_result(plan: Plan, res: OrCallResult, keep) -> dict:
    obs = plan.observations
    fits = []
    for r in range(obs - 1):
        f = keep["fits"][r]
        fits.append({
            "node": f.node, "kind": KIND_NAMES[f.kind], "exercise": bool(f.exercise),
            "estimated_continuation": f.estimated_continuation, "itm_paths": f.itm_paths,
            "required_paths": f.required_paths, "model": model_dict(f.model),
            "near_ties": f.near_ties,
        })
    return {
        "price": res.price, "paths": res.paths,
        "sd": res.sd if res.has_sd else None, "se": res.se if res.has_sd else None,
        "observation_nodes": keep["nodes"].tolist(),
        "exercise_counts": keep["counts"].tolist(), "fits": fits,
        "near_tie_training_decisions": int(res.near_tie_training),
        "near_tie_evaluation_decisions": int(res.near_tie_evaluation),
    }


# Default window of the tie counts (the GPU executor's default: exact ties).
TIE_WINDOW = 0.0


def _estimator(strike, degree, rank_tolerance, condition_limit, tie_rule, support,
               tie_window=TIE_WINDOW):
    return OrEstimator(float(strike), int(degree), float(rank_tolerance), float(condition_limit),
                       int(tie_rule), int(support), float(tie_window))


def price_lsm_call(plan: Plan, initial_spot: float, strike: float, training_paths: int,
                   evaluation_paths: int, training_seed: int, training_domain: int,
                   evaluation_seed: int, evaluation_domain: int, degree: int = 4,
                   rank_tolerance: float = 1e-12, condition_limit: float = 1e8,
                   tie_rule: int = 0, insufficient_support: int = 0,
                   tie_window: float = TIE_WINDOW) -> dict:
    s, keep_plan = plan.struct()
    obs = plan.observations
    keep = {"nodes": np.zeros(obs, np.uint32), "counts": np.zeros(obs, np.uint32),
            "fits": (OrNodeFit * max(obs, 1))()}
    res = OrCallResult(0, 0, 0, 0, 0, obs, _p(keep["nodes"], C.c_uint32),
                       _p(keep["counts"], C.c_uint32), keep["fits"], 0)
    est = _estimator(strike, degree, rank_tolerance, condition_limit, tie_rule,
                     insufficient_support, tie_window)
    err = OrError()
    rc = lib().or_price_lsm_call(C.byref(s), C.c_double(initial_spot), C.byref(est),
                                 C.c_uint32(training_paths), C.c_uint32(evaluation_paths),
                                 C.c_uint64(training_seed), C.c_uint32(training_domain),
                                 C.c_uint64(evaluation_seed), C.c_uint32(evaluation_domain),
                                 C.byref(res), C.byref(err))
    if rc:
        raise OracleError(err)
    return _call_result(plan, res, keep)


def call_pathwise_greeks(plan: Plan, initial_spot: float, strike: float, training_paths: int,
                         evaluation_paths: int, training_seed: int, training_domain: int,
                         evaluation_seed: int, evaluation_domain: int, degree: int = 4,
                         rank_tolerance: float = 1e-12, condition_limit: float = 1e8,
                         tie_rule: int = 0, insufficient_support: int = 0) -> dict:
    """Frozen-policy pathwise delta, vega, rho and dividend-yield sensitivity of one call."""
    s, keep_plan = plan.struct()
    est = _estimator(strike, degree, rank_tolerance, condition_limit, tie_rule,
                     insufficient_support)
    out = np.zeros(4)
    err = OrError()
    rc = lib().or_call_pathwise_greeks(C.byref(s), C.c_double(initial_spot), C.byref(est),
                                       C.c_uint32(training_paths), C.c_uint32(evaluation_paths),
                                       C.c_uint64(training_seed), C.c_uint32(training_domain),
                                       C.c_uint64(evaluation_seed), C.c_uint32(evaluation_domain),
                                       _p(out, C.c_double), C.byref(err))
    if rc:
        raise OracleError(err)
    return dict(zip(("delta", "vega", "rho", "dividend_yield"), out.tolist()))


def fit_evaluate_paths(plan: Plan, initial_spot: float, strike: float, training: np.ndarray,
                       evaluation: np.ndarray, degree: int = 4, rank_tolerance: float = 1e-12,
                       condition_limit: float = 1e8, tie_rule: int = 0,
                       insufficient_support: int = 0, tie_window: float = TIE_WINDOW) -> dict:
    s, keep_plan = plan.struct()
    obs = plan.observations
    tr = np.ascontiguousarray(training, dtype=np.float64)
    ev = np.ascontiguousarray(evaluation, dtype=np.float64)
    keep = {"nodes": np.zeros(obs, np.uint32), "counts": np.zeros(obs, np.uint32),
            "fits": (OrNodeFit * max(obs, 1))()}
    res = OrCallResult(0, 0, 0, 0, 0, obs, _p(keep["nodes"], C.c_uint32),
                       _p(keep["counts"], C.c_uint32), keep["fits"], 0)
    est = _estimator(strike, degree, rank_tolerance, condition_limit, tie_rule,
                     insufficient_support, tie_window)
    err = OrError()
    rc = lib().or_fit_evaluate_paths(C.byref(s), C.c_double(initial_spot), C.byref(est),
                                     _p(tr, C.c_double), C.c_uint32(tr.shape[1]),
                                     _p(ev, C.c_double), C.c_uint32(ev.shape[1]), C.byref(res),
                                     C.byref(err))
    if rc:
        raise OracleError(err)
    return _call_result(plan, res, keep)


# ---------------------------------------------------------------------------------- grouped

GROUPED_COLUMNS = (
    ("prices", np.float64), ("sample_standard_deviations", np.float64),
    ("standard_errors", np.float64), ("origin_decisions_fitted", np.bool_),
    ("origin_exercised", np.bool_), ("regression_nodes", np.uint32),
    ("no_training_support_nodes", np.uint32), ("insufficient_exercise_support_nodes", np.uint32),
    ("minimum_in_the_money_training_paths", np.uint32),
    ("maximum_regression_condition_numbers", np.float64), ("early_exercise_paths", np.uint32),
    ("expiry_paths", np.uint32),
)


@dataclass
class GroupedRequest:
    """Flattened group-addressed strike request (the C-ABI request's field meaning)."""

    group_offsets: np.ndarray
    initial_spots: np.ndarray
    strikes: np.ndarray
    maturities: np.ndarray
    rates: np.ndarray
    volatilities: np.ndarray
    training_paths: int
    evaluation_paths: int
    training_seed: int
    evaluation_seed: int
    training_domain_start: int
    evaluation_domain_start: int
    dividend_yields: np.ndarray | None = None
    cash_offsets: np.ndarray | None = None
    cash_times: np.ndarray | None = None
    cash_amounts: np.ndarray | None = None
    cash_policy: int = 0
    mesh_kind: int = 0
    base_intervals: int = 1
    origin_exercise: int = 0
    exercise_offsets: np.ndarray | None = None
    exercise_times: np.ndarray | None = None
    degree: int = 4
    rank_tolerance: float = 1e-12
    condition_limit: float = 1e8
    tie_rule: int = 0
    insufficient_support: int = 0
    training_path_counts: np.ndarray | None = None
    evaluation_path_counts: np.ndarray | None = None
    # escrowed cash-dividend model for the cash columns: None (spot cash model), 0 prepaid-forward
    # volatility, 1 spot volatility matched at the origin
    escrowed: int | None = None
    # window of the near-tie counts
    tie_window: float = TIE_WINDOW

    @property
    def dividend_model(self) -> int:
        if self.cash_offsets is None:
            return 0
        return 1 if self.escrowed is None else 2

    def struct(self) -> tuple[OrGroupedInput, list]:
        f64 = lambda a: np.ascontiguousarray(a if a is not None else [0.0], dtype=np.float64)
        i64 = lambda a: np.ascontiguousarray(a if a is not None else [0, 0], dtype=np.int64)
        keep = [np.ascontiguousarray(self.group_offsets, dtype=np.uint32),
                f64(self.initial_spots), f64(self.maturities), f64(self.rates),
                f64(self.volatilities), f64(self.strikes), f64(self.dividend_yields),
                i64(self.cash_offsets), f64(self.cash_times), f64(self.cash_amounts),
                i64(self.exercise_offsets), f64(self.exercise_times)]
        counts = [None if c is None else np.ascontiguousarray(c, dtype=np.uint32)
                  for c in (self.training_path_counts, self.evaluation_path_counts)]
        keep += [c for c in counts if c is not None]
        cptr = [C.cast(None, C.POINTER(C.c_uint32)) if c is None else _p(c, C.c_uint32)
                for c in counts]
        groups = len(self.initial_spots)
        s = OrGroupedInput(
            groups, _p(keep[0], C.c_uint32), _p(keep[1], C.c_double), _p(keep[2], C.c_double),
            _p(keep[3], C.c_double), _p(keep[4], C.c_double), len(self.strikes),
            _p(keep[5], C.c_double), self.dividend_model,
            _p(keep[6], C.c_double), _p(keep[7], C.c_int64), _p(keep[8], C.c_double),
            _p(keep[9], C.c_double), int(self.cash_policy), int(self.mesh_kind),
            int(self.base_intervals), int(self.origin_exercise), _p(keep[10], C.c_int64),
            _p(keep[11], C.c_double), int(self.training_paths), int(self.evaluation_paths),
            cptr[0], cptr[1], int(self.training_seed), int(self.evaluation_seed), int(self.training_domain_start),
            int(self.evaluation_domain_start), int(self.degree), float(self.rank_tolerance),
            float(self.condition_limit), int(self.tie_rule), int(self.insufficient_support),
            int(self.escrowed or 0), float(self.tie_window))
        return s, keep


def price_grouped(req: GroupedRequest, threads: int = 0) -> dict[str, np.ndarray]:
    """The 12 row columns (and tie counts). ``threads`` <= 1 prices groups in order on one thread;
    more threads price groups in parallel, each group serially (the reference's threading)."""
    s, keep = req.struct()
    rows = len(req.strikes)
    cols = {name: np.zeros(rows, dtype=(np.uint8 if dt is np.bool_ else dt))
            for name, dt in GROUPED_COLUMNS}
    ctype = {np.float64: C.c_double, np.uint32: C.c_uint32, np.bool_: C.c_uint8}
    ptrs = [_p(cols[name], ctype[dt]) for name, dt in GROUPED_COLUMNS]
    ties = {name: np.zeros(rows, dtype=np.uint64)
            for name in ("near_tie_training_decisions", "near_tie_evaluation_decisions")}
    ptrs += [_p(v, C.c_uint64) for v in ties.values()]
    out = OrBatchOutput(*ptrs)
    err = OrError()
    rc = lib().or_price_grouped(C.byref(s), C.byref(out), int(threads), C.byref(err))
    if rc:
        raise OracleError(err)
    for name, dt in GROUPED_COLUMNS:
        if dt is np.bool_:
            cols[name] = cols[name].astype(np.bool_)
    cols.update(ties)
    return cols


def group_grid(req: GroupedRequest, group: int) -> dict:
    s, keep = req.struct()
    cap = 4096 + (req.base_intervals if req.mesh_kind == 0 else 0)
    if req.exercise_times is not None:
        cap += len(req.exercise_times)
    if req.cash_times is not None:
        cap += len(req.cash_times)
    times = np.zeros(cap)
    ex = np.zeros(cap, np.uint32)
    dn = np.zeros(cap, np.uint32)
    da = np.zeros(cap)
    ne = C.c_uint32()
    nd = C.c_uint32()
    n = lib().or_group_grid(C.byref(s), C.c_uint32(group), _p(times, C.c_double), C.c_uint32(cap),
                            _p(ex, C.c_uint32), C.byref(ne), _p(dn, C.c_uint32),
                            _p(da, C.c_double), C.byref(nd))
    return {"times": times[:n].copy(), "exercise_nodes": ex[:ne.value].copy(),
            "dividend_nodes": dn[:nd.value].copy(), "dividend_amounts": da[:nd.value].copy()}


def grouped_models(req: GroupedRequest, group: int, rows_per_strike: int) -> list[list[dict]]:
    s, keep = req.struct()
    a, b = int(req.group_offsets[group]), int(req.group_offsets[group + 1])
    buf = (OrNodeFit * ((b - a) * rows_per_strike))()
    err = OrError()
    rc = lib().or_grouped_models(C.byref(s), C.c_uint32(group), buf, C.c_uint32(rows_per_strike),
                                 C.byref(err))
    if rc:
        raise OracleError(err)
    out = []
    for k in range(b - a):
        row = []
        for r in range(rows_per_strike):
            f = buf[k * rows_per_strike + r]
            row.append({"node": f.node, "kind": KIND_NAMES[f.kind], "exercise": bool(f.exercise),
                        "estimated_continuation": f.estimated_continuation,
                        "itm_paths": f.itm_paths, "model": model_dict(f.model)})
        out.append(row)
    return out


def grouped_pathwise_greeks(req: GroupedRequest, group: int) -> dict[str, np.ndarray]:
    s, keep = req.struct()
    a, b = int(req.group_offsets[group]), int(req.group_offsets[group + 1])
    res = {k: np.zeros(b - a) for k in ("delta", "vega", "rho", "dividend_yield")}
    err = OrError()
    rc = lib().or_grouped_pathwise_greeks(
        C.byref(s), C.c_uint32(group), _p(res["delta"], C.c_double), _p(res["vega"], C.c_double),
        _p(res["rho"], C.c_double), _p(res["dividend_yield"], C.c_double), C.byref(err))
    if rc:
        raise OracleError(err)
    return res


def grouped_frozen_prices(fit_req: GroupedRequest, eval_req: GroupedRequest, group: int, *,
                          stops: np.ndarray | None = None, record: bool = False) -> dict:
    """Prices of ``eval_req``'s group under the policy fitted on ``fit_req``'s group. With
    ``record`` the stopping rows are returned; passing ``stops`` holds them fixed."""
    fs, fkeep = fit_req.struct()
    es, ekeep = eval_req.struct()
    a, b = int(fit_req.group_offsets[group]), int(fit_req.group_offsets[group + 1])
    ne = int(eval_req.evaluation_path_counts[group]) if eval_req.evaluation_path_counts is not None \
        else int(eval_req.evaluation_paths)
    prices = np.zeros(b - a)
    mode = 2 if stops is not None else (1 if record else 0)
    buf = (np.ascontiguousarray(stops, dtype=np.int32) if stops is not None
           else np.zeros((b - a) * ne if record else 1, dtype=np.int32))
    err = OrError()
    rc = lib().or_grouped_frozen_prices(C.byref(fs), C.byref(es), C.c_uint32(group),
                                        _p(buf, C.c_int32), C.c_int(mode),
                                        _p(prices, C.c_double), C.byref(err))
    if rc:
        raise OracleError(err)
    return {"prices": prices, "stops": buf.reshape(b - a, ne) if mode else None}


def grouped_paths(req: GroupedRequest, group: int, *, evaluation: bool = False) -> np.ndarray:
    """Observation matrix [obs, paths] of one group's training or evaluation population, as the
    fit and the evaluation read it (escrowed groups: the reconstructed stock)."""
    s, keep = req.struct()
    counts = req.evaluation_path_counts if evaluation else req.training_path_counts
    n = int(counts[group]) if counts is not None else int(
        req.evaluation_paths if evaluation else req.training_paths)
    obs = len(group_grid(req, group)["exercise_nodes"]) + 1
    out = np.zeros((obs, n))
    err = OrError()
    rc = lib().or_grouped_paths(C.byref(s), C.c_uint32(group), C.c_int(int(evaluation)),
                                _p(out, C.c_double), C.byref(err))
    if rc:
        raise OracleError(err)
    return out


def escrowed_prepaid_cash(req: GroupedRequest, group: int) -> np.ndarray:
    """Prepaid cash per grid node of an escrowed group (index 0: origin value)."""
    s, keep = req.struct()
    cap = len(group_grid(req, group)["times"])
    out = np.zeros(cap)
    err = OrError()
    n = lib().or_escrowed_prepaid_cash(C.byref(s), C.c_uint32(group), _p(out, C.c_double),
                                       C.c_uint32(cap), C.byref(err))
    if n < 0:
        raise OracleError(err)
    return out[:n]


def set_threads(threads: int) -> None:
    """Threads of a single call's path-parallel simulation."""
    lib().or_set_threads(int(threads))


def threads_available() -> int:
    return len(os.sched_getaffinity(0))
