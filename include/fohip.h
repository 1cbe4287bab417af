/*
 * fohip: C ABI of the MI355X-native Longstaff-Schwartz engine (libfohip.so).
 *
 * Plain C types only (pointers + sizes, structs of pointers); no torch or HIP types cross this
 * boundary. Every entry point returns 0 on success or a FOHIP_E_* code and fills `fohip_error`.
 * Failures are atomic: no partial result is reported as success, and no other precision, device or
 * CPU path is substituted.
 *
 * Reference interfaces replaced (jorgensandhaug/finoptions-engine):
 *   fohip_executor_create/destroy  finoptions-runtime CpuExecutor::try_new
 *                                  (crates/finoptions-runtime/src/lib.rs:341-389); Python
 *                                  CpuExecutor (python/finoptions/simulation.py:40-62)
 *   fohip_price_lsm_call           _native.price_lsm_call_seeded_f64
 *                                  (crates/finoptions-python/src/lib.rs:1170-1248) ->
 *                                  LsmCallEstimator::price_seeded (crates/finoptions-lsm/src/policy.rs:306)
 *   fohip_price_lsm_grouped        _native.price_lsm_call_grouped_batch_{continuous_yield,
 *                                  cash_dividends,escrowed_cash_dividends}_f64
 *                                  (crates/finoptions-python/src/lib.rs:1885-2061) and the row
 *                                  batches _native.price_lsm_call_batch_*_f64 ->
 *                                  GroupedBatchLsmCallInputs::price (batch/pricing.rs:422),
 *                                  BatchLsmCallInputs::price (batch/pricing.rs:191)
 *   fohip_simulate_gbm_seeded      _native.simulate_gbm_seeded_f64 (lib.rs:1020-1060) ->
 *                                  GbmPathPlan::simulate_seeded_with (crates/finoptions-sim/src/executor.rs:34)
 *   fohip_simulate_gbm_seeded_device  the same export into caller device memory (no host copy)
 *   fohip_philox_blocks            philox4x32_10 (crates/finoptions-sim/src/random/philox.rs:32) over
 *                                  PhiloxPathStream counters (philox.rs:64-86); parity probe
 *   fohip_stream_normals           sample_standard_normal (crates/finoptions-sim/src/random/normal.rs:96);
 *                                  parity probe
 *   fohip_validate_*               _native.validate_event_schedule / validate_log_discount_curve /
 *                                  validate_cumulative_dividend_yield_curve / validate_constant_volatility
 *                                  / validate_cash_dividend_amount (lib.rs:407-470, 703-722)
 */
#ifndef FOHIP_H
#define FOHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define FOHIP_ABI_VERSION 5

/* ---- error codes (the numbering is shared with the CPU oracle) ---- */
enum {
    FOHIP_OK = 0,
    /* 1xx: financial / request validation (reference InputError) */
    FOHIP_E_INVALID_ARGUMENT = 100,
    FOHIP_E_INVALID_INITIAL_SPOT = 101,
    FOHIP_E_INVALID_STRIKE = 102,
    FOHIP_E_INVALID_VOLATILITY = 103,
    FOHIP_E_INVALID_TIME_GRID = 104,
    FOHIP_E_INVALID_LOG_DISCOUNT_CURVE = 105,
    FOHIP_E_INVALID_YIELD_CURVE = 106,
    FOHIP_E_INVALID_SCHEDULE = 107,
    FOHIP_E_INVALID_CASH_DIVIDEND = 108,
    FOHIP_E_TIME_GRID_MISMATCH = 109,
    FOHIP_E_INVALID_PATH_COUNT = 110,
    FOHIP_E_INVALID_REGRESSION_SPEC = 111,
    FOHIP_E_UNSUPPORTED_DEGREE = 112,
    /* 2xx: simulation (reference SimulationError) */
    FOHIP_E_NON_FINITE_TRANSITION = 200,
    FOHIP_E_NORMAL_REJECTION_LIMIT = 201,
    FOHIP_E_PATH_NAMESPACE_OVERFLOW = 202,
    FOHIP_E_EMPTY_PATH_SET = 203,
    /* 3xx: regression at an exercise node (reference RegressionError) */
    FOHIP_E_INSUFFICIENT_OBSERVATIONS = 300,
    FOHIP_E_DEGENERATE_PREDICTOR = 301,
    FOHIP_E_RANK_DEFICIENT = 302,
    FOHIP_E_CONDITION_LIMIT = 303,
    FOHIP_E_NON_FINITE_FIT = 304,
    FOHIP_E_JACOBI_NOT_CONVERGED = 305,
    FOHIP_E_NON_FINITE_PREDICTOR = 306,
    FOHIP_E_NON_FINITE_RESPONSE = 307,
    FOHIP_E_NON_FINITE_PREDICTION = 308,
    /* 4xx: policy fitting / evaluation (reference LsmError) */
    FOHIP_E_RANDOM_DOMAIN_COLLISION = 400,
    FOHIP_E_NON_FINITE_DISCOUNTED_PAYOFF = 401,
    FOHIP_E_OUTSIDE_TRAINING_SUPPORT = 402,
    FOHIP_E_UNREPRESENTABLE_DISCOUNT = 403,
    FOHIP_E_PATH_COUNT_OUTSIDE_DOMAIN = 404,
    FOHIP_E_STOPPING_COUNT_MISMATCH = 405,
    /* 5xx: batch request validation (reference BatchLsmError -> InputError) */
    FOHIP_E_EMPTY_BATCH = 500,
    FOHIP_E_ROW_COUNT_MISMATCH = 501,
    FOHIP_E_GROUP_OFFSETS = 502,
    FOHIP_E_INVALID_ROW_VALUE = 503,
    FOHIP_E_RANDOM_DOMAIN_RANGE = 504,
    FOHIP_E_CASH_SCHEDULE = 505,
    FOHIP_E_EXERCISE_MESH = 506,
    FOHIP_E_INSUFFICIENT_EVALUATION_PATHS = 507,
    FOHIP_E_ESCROWED_PREPAID_FORWARD = 508,   /* nonpositive F0, non-finite prepaid cash, or an
                                                 invalid spot-matched volatility */
    /* 9xx: execution (reference ExecutionError) */
    FOHIP_E_NO_DEVICE = 900,
    FOHIP_E_HIP = 901,
    FOHIP_E_OUT_OF_MEMORY = 902,
    FOHIP_E_UNSUPPORTED = 903,
    FOHIP_E_INTERNAL = 904
};

/* phase of a failure inside one pricing request */
enum {
    FOHIP_PHASE_NONE = 0,
    FOHIP_PHASE_TRAINING_SIMULATION = 1,
    FOHIP_PHASE_FIT = 2,
    FOHIP_PHASE_EVALUATION_SIMULATION = 3,
    FOHIP_PHASE_EVALUATION = 4
};

typedef struct {
    int32_t code;
    int32_t phase;
    int64_t row;       /* strike row of a batch, or -1 */
    int64_t node;      /* exercise / transition node, or -1 */
    int64_t path;      /* path index, or -1 */
    int64_t index;     /* code-specific integer (interval, rank, observed count, ...) */
    double observed;   /* code-specific value */
    double limit;      /* code-specific value */
    char message[512];
} fohip_error;

/* ---- executor ---- */
/* F64: the reference's algorithm and arithmetic (Welford pass + d + 1 Stieltjes passes per node).
 * FUSED_F64: one fused HBM pass per node (pivoted power sums, moment-form Stieltjes solve), binary64
 *   spots and cashflows, binary64 exercise decisions; prices within 1e-10 of F64 in the absence of
 *   near-tie decisions, Greek tangents accumulated in binary32.
 * F32_BULK: FUSED_F64 with binary32 storage of spots and cashflows. Binary32 decision inputs flip
 *   rare near-tie exercise decisions, which backward induction amplifies: typically below 0.01
 *   SE, occasionally 0.1-0.5 SE on single strikes. */
enum { FOHIP_PRECISION_F64 = 0, FOHIP_PRECISION_F32_BULK = 1, FOHIP_PRECISION_FUSED_F64 = 2 };

typedef struct fohip_executor fohip_executor;

/* Binds one HIP device explicitly. Fails with FOHIP_E_NO_DEVICE when the device does not exist;
 * there is no CPU fallback. `precision` is one of FOHIP_PRECISION_*. */
int fohip_executor_create(int32_t device, int32_t precision, fohip_executor **out,
                          fohip_error *err);
int fohip_executor_destroy(fohip_executor *executor);
int fohip_executor_info(const fohip_executor *executor, int32_t *device, int32_t *precision,
                        char *device_name, size_t name_capacity);
int fohip_device_count(int32_t *count);
int fohip_abi_version(void);

/* Launch parameters an autotuner may change. Results are bitwise invariant to them: reduction
 * trees are fixed by the path chunks, not by these values. Zero fields keep defaults. */
enum { FOHIP_CASHFLOWS_AUTO = 0, FOHIP_CASHFLOWS_VALUE_COLUMNS = 1, FOHIP_CASHFLOWS_STOPPING_ROWS = 2 };
enum { FOHIP_BATCHES_SERIAL = 0, FOHIP_BATCHES_OVERLAPPED = 1 };
enum { FOHIP_FIT_AUTO = 0, FOHIP_FIT_PER_NODE = 1, FOHIP_FIT_SWEEP = 2 };
typedef struct {
    uint32_t eval_strike_tile;      /* strikes per evaluation launch column (1..16; Greeks 1, 4, 6, 8) */
    /* Fused methods: how the backward induction keeps each (strike, path) cashflow. Value
     * columns hold it (8 B, binary32 for FP32 bulk); stopping rows hold one byte, the path's
     * stopping row, and rebuild the cashflow from the group's spot matrix (observation rows
     * <= 255): 8x less cashflow memory and traffic, bitwise the same FUSED_F64 results, but
     * slower on the shapes measured. Auto keeps value columns. */
    uint32_t fit_cashflows;
    /* How the memory-bounded batches of one request share the device. Serial (the default)
     * runs them one after another, each within a quarter of device memory. Overlapped keeps two
     * batches in flight in two slots of an eighth each: one batch's backward induction, bound by
     * HBM bandwidth, runs on the executor's stream while the next batch's simulation and the
     * previous batch's evaluation, bound by vector ALU issue, run on a second stream. On the
     * measured shapes the overlap gains less than the halved batches cost: the generation
     * kernels fill every register slot, so the fit's blocks rarely run beside them. */
    uint32_t batch_schedule;
    /* Bytes of paths and cashflows per batch slot, in MiB (0: a quarter of device memory under
     * the serial schedule, an eighth per slot when overlapped). Smaller slots mean more, shorter
     * batches; results do not depend on the batch split. */
    uint32_t batch_mib;
    /* Fused methods: how the backward induction is launched. Per-node launches run one fit pass
     * and one reduce per exercise row over every group of a batch. The sweep, for requests whose
     * groups all hold one strike, runs the whole backward induction of a batch in one persistent
     * launch: each group's paths are split over a team of blocks that keep their spot rows and
     * cashflows on chip from the last row to the first and meet at a device-memory counter per
     * row, so each row is read from HBM once. Both give bitwise the same results. Auto is the
     * per-node launches: on the shapes measured the sweep's per-row gathering of a team costs
     * more than the HBM traffic it saves. */
    uint32_t fit_schedule;
} fohip_tuning;
int fohip_set_tuning(fohip_executor *executor, const fohip_tuning *tuning);

/* Window of the tie counts reported with every result. The default 0 counts exact ties,
 * intrinsic == continuation: the decisions the tie rule settles, which the other rule would
 * decide the other way. A window w in (0, 1] counts every |intrinsic - continuation| <=
 * w |continuation| instead (e.g. 2^-40, wider than the rounding differences between the methods'
 * continuations); it is a diagnostic mode whose evaluation runs single-strike tiles and is
 * several times slower. Results other than the counts do not depend on the window. No
 * counterpart in the reference. */
int fohip_set_near_tie_window(fohip_executor *executor, double window);

/* Per-kernel timing with HIP events recorded on the executor's stream around every launch
 * (off by default; it adds one event pair per launch). Statistics accumulate until reset. */
typedef struct {
    char name[48];
    uint64_t launches;
    double total_ms;
} fohip_kernel_stat;
int fohip_set_profiling(fohip_executor *executor, int32_t enabled);
int fohip_get_profile(fohip_executor *executor, fohip_kernel_stat *out, size_t capacity,
                      size_t *count);
int fohip_reset_profile(fohip_executor *executor);

/* ---- Greeks (pathwise, frozen stopping policy) ---- */
enum {
    FOHIP_GREEK_DELTA = 1u << 0,
    FOHIP_GREEK_VEGA = 1u << 1,
    FOHIP_GREEK_RHO = 1u << 2,
    FOHIP_GREEK_DIVIDEND_YIELD = 1u << 3
};

/* ---- single scheduled call: price_lsm_call ---- */
enum { FOHIP_DIVIDEND_LIQUIDATOR = 0, FOHIP_DIVIDEND_SURVIVOR = 1 };
enum { FOHIP_TIE_STRICTLY_GREATER = 0, FOHIP_TIE_GREATER_OR_EQUAL = 1 };
enum { FOHIP_SUPPORT_FAIL = 0, FOHIP_SUPPORT_CONTINUE = 1 };

typedef struct {
    double initial_spot;
    double strike;
    const double *schedule_times;       size_t schedule_nodes;
    const double *curve_times;          const double *log_discounts;      size_t curve_nodes;
    const double *yield_times;          const double *cumulative_yields;  size_t yield_nodes;
    const uint32_t *dividend_nodes;     const double *dividend_amounts;   size_t dividends;
    const uint32_t *exercise_nodes;     size_t exercise_count;
    uint32_t expiry_node;
    double volatility;
    int32_t dividend_policy;
    uint64_t training_paths, evaluation_paths;
    uint64_t training_seed;   uint32_t training_domain;
    uint64_t evaluation_seed; uint32_t evaluation_domain;
    uint32_t polynomial_degree;
    double rank_tolerance, condition_limit;
    int32_t tie_rule;
    int32_t insufficient_support;
    uint32_t greeks;                    /* FOHIP_GREEK_* mask; 0 = price only */
} fohip_call_request;

/* one exercise node's fit (kind: 0 origin decision, 1 no ITM paths, 2 insufficient support,
 * 3 regression) -- the 11 fields of NativeExerciseDiagnostic (lib.rs:1060-1072) */
typedef struct {
    uint32_t node;
    int32_t kind;
    int32_t exercise;
    double estimated_continuation;
    uint32_t in_the_money_paths;
    uint32_t observations;
    uint32_t columns;
    uint32_t rank;
    double condition_number;
    double predictor_mean;
    double predictor_scale;
    /* fitted orthogonal recurrence (not part of the reference tuple; parity evidence) */
    double alpha[4], beta[4], coefficients[5];
} fohip_node_diagnostic;

typedef struct {
    double price;
    uint32_t paths;
    int32_t has_standard_error;         /* 0 when paths == 1 (reference returns None) */
    double sample_standard_deviation;
    double standard_error;
    uint32_t observation_rows;          /* = exercise_count + 1 */
    uint32_t *observation_nodes;        /* [exercise_count + 1], caller buffer */
    uint32_t *exercise_counts;          /* [exercise_count + 1], caller buffer */
    fohip_node_diagnostic *diagnostics; /* [exercise_count], caller buffer */
    double greeks[4];                   /* delta, vega, rho, dividend_yield; NaN when not requested */
    double greek_standard_errors[4];
    /* Tied rule applications (see fohip_set_near_tie_window), training (backward pass) and
     * evaluation: by default the exact ties, whose outcome the tie rule settles. No counterpart in
     * the reference; reported so that decision flips at ties are accounted for. */
    uint64_t near_tie_training_decisions;
    uint64_t near_tie_evaluation_decisions;
} fohip_call_result;

int fohip_price_lsm_call(fohip_executor *executor, const fohip_call_request *request,
                         fohip_call_result *result, fohip_error *err);

/* Allocation-free variant of fohip_price_lsm_call: the same request, result and errors, priced
 * with bitwise the same results. No counterpart in the reference; it is for callers that cannot
 * allocate (code under HIP graph capture, tight serving loops, harnesses that forbid allocation).
 * Every device buffer of the pricing is carved from one fixed region of
 * FOHIP_NO_ALLOC_SCRATCH_BYTES per device (a file-scope device array, held by each device the
 * library runs on until the process ends) and host transfers are staged through pageable memory,
 * so the call makes no hipMalloc, hipFree or pinned-host allocation. A request whose buffers do
 * not fit the region fails with FOHIP_E_UNSUPPORTED before any work is queued (about 2.5 M
 * training + evaluation paths for a call with one exercise date). All calls on one device share
 * the region and it keeps the last call's contents: calls that overlapped on one device would
 * race, so the library serializes them per device within a process. Per-kernel profiling must be
 * off (it creates events). */
#define FOHIP_NO_ALLOC_SCRATCH_BYTES (64u << 20)
int fohip_price_lsm_call_noalloc(fohip_executor *executor, const fohip_call_request *request,
                                 fohip_call_result *result, fohip_error *err);

/* ---- group-addressed strikes: price_lsm_call_grouped_batch ---- */
enum { FOHIP_DIVIDENDS_CONTINUOUS_YIELD = 0, FOHIP_DIVIDENDS_CASH = 1, FOHIP_DIVIDENDS_ESCROWED_CASH = 2 };
/* escrowed cash dividends (EscrowedVolatilityConvention, batch_lsm.py:52-56): the volatility
 * column is the prepaid forward's, or a spot volatility rescaled by S0 / F0 */
enum { FOHIP_ESCROWED_PREPAID_FORWARD = 0, FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN = 1 };
enum { FOHIP_MESH_UNIFORM = 0, FOHIP_MESH_CASH_DIVIDEND_DATES = 1, FOHIP_MESH_EXPLICIT = 2 };
enum { FOHIP_ORIGIN_EXCLUDED = 0, FOHIP_ORIGIN_INCLUDED = 1 };
enum { FOHIP_BOUNDARY_STRICTLY_BEFORE = 0, FOHIP_BOUNDARY_AT_OR_BEFORE = 1 };

typedef struct {
    const uint32_t *group_offsets;      /* [groups + 1], uint32 strike-row boundaries */
    size_t groups;
    const double *initial_spots;        /* [groups] */
    const double *maturities;           /* [groups] */
    const double *rates;                /* [groups] constant continuously compounded */
    const double *volatilities;         /* [groups] */
    const double *strikes;              /* [rows] */
    size_t rows;
    int32_t dividend_model;             /* FOHIP_DIVIDENDS_* */
    const double *dividend_yields;      /* [groups] (continuous yield) */
    const int64_t *cash_offsets;        /* [groups + 1] Arrow-style offsets (cash and escrowed) */
    const double *cash_times;
    const double *cash_amounts;
    size_t cash_events;                 /* length of cash_times / cash_amounts */
    int32_t cash_policy;                /* FOHIP_DIVIDEND_* (cash model) */
    int32_t maturity_boundary;          /* FOHIP_BOUNDARY_* */
    int32_t mesh_kind;                  /* FOHIP_MESH_* */
    uint32_t base_intervals;            /* uniform mesh */
    int32_t origin_exercise;            /* FOHIP_ORIGIN_* (uniform and explicit meshes) */
    const int64_t *exercise_offsets;    /* [groups + 1] (explicit mesh) */
    const double *exercise_times;
    size_t exercise_events;
    uint64_t training_paths, evaluation_paths;
    /* Optional per-group path counts (ragged batches; no counterpart in the reference, whose Lsm
     * method carries one count for every row). Null pointers use training_paths/evaluation_paths. */
    const uint32_t *training_path_counts;    /* [groups] or null */
    const uint32_t *evaluation_path_counts;  /* [groups] or null */
    uint64_t training_seed, evaluation_seed;
    uint32_t training_domain_start, evaluation_domain_start;
    uint32_t polynomial_degree;
    double rank_tolerance, condition_limit;
    int32_t tie_rule;
    int32_t insufficient_support;
    uint32_t greeks;                    /* FOHIP_GREEK_* mask */
    int32_t escrowed_volatility;        /* FOHIP_ESCROWED_* (escrowed model) */
} fohip_grouped_request;

/* The 12 row-ordered columns of NativeBatchLsmResult (lib.rs:1417-1430) plus optional Greek
 * columns (null pointers skip them). All caller-owned, length `rows`. */
typedef struct {
    double *prices;
    double *sample_standard_deviations;
    double *standard_errors;
    uint8_t *origin_decisions_fitted;
    uint8_t *origin_exercised;
    uint32_t *regression_nodes;
    uint32_t *no_training_support_nodes;
    uint32_t *insufficient_exercise_support_nodes;
    uint32_t *minimum_in_the_money_training_paths;
    double *maximum_regression_condition_numbers;
    uint32_t *early_exercise_paths;
    uint32_t *expiry_paths;
    double *delta, *vega, *rho, *dividend_yield;
    double *delta_se, *vega_se, *rho_se, *dividend_yield_se;
    /* near-tie rule applications per row, training and evaluation (see fohip_call_result) */
    uint64_t *near_tie_training_decisions, *near_tie_evaluation_decisions;
} fohip_batch_result;

int fohip_price_lsm_grouped(fohip_executor *executor, const fohip_grouped_request *request,
                            fohip_batch_result *result, fohip_error *err);

/* Fitted rules of every (strike row, exercise row) of one group of `request` (diagnostic and
 * parity evidence; the group is priced again with identical streams). `out` receives
 * strikes x exercise_rows entries, strike-major; `exercise_rows` receives the row count. */
int fohip_grouped_diagnostics(fohip_executor *executor, const fohip_grouped_request *request,
                              uint32_t group, fohip_node_diagnostic *out, size_t capacity,
                              uint32_t *exercise_rows, fohip_error *err);

/* ---- path export and random-stream parity probes ---- */
typedef struct {
    const double *times;               size_t nodes;
    const double *log_discounts;
    const double *cumulative_yields;
    const uint32_t *dividend_nodes;    const double *dividend_amounts;   size_t dividends;
    const uint32_t *exercise_nodes;    size_t exercise_count;
    uint32_t expiry_node;
    double volatility;
    int32_t dividend_policy;
} fohip_path_plan;

/* Time-major [exercise_count + 1, paths] states, written to a host buffer `out` (FP64) or, for
 * an FP32-bulk executor, converted from FP32 paths. */
int fohip_simulate_gbm_seeded(fohip_executor *executor, const fohip_path_plan *plan,
                              double initial_spot, uint64_t paths, uint64_t seed, uint32_t domain,
                              double *out, fohip_error *err);

/* The same time-major states written straight into caller device memory on the executor's device
 * (simulate_gbm_seeded, python/finoptions/simulation.py:262-332, with the observation matrix left
 * in HBM for torch/DLPack consumers instead of copied to a host array). `device_out` holds
 * (exercise_count + 1) * paths elements of `dtype` (binary32 rounds each binary64 state once).
 * The simulation is ordered after the work already queued on `stream` (a hipStream_t, or null)
 * and the call returns once the states are complete. */
enum { FOHIP_DTYPE_F64 = 0, FOHIP_DTYPE_F32 = 1 };
int fohip_simulate_gbm_seeded_device(fohip_executor *executor, const fohip_path_plan *plan,
                                     double initial_spot, uint64_t paths, uint64_t seed,
                                     uint32_t domain, void *device_out, int32_t dtype, void *stream,
                                     fohip_error *err);

/* Raw Philox4x32-10 output words: out[(p * blocks + b) * 4 + w] for path path0 + p, block b. */
int fohip_philox_blocks(fohip_executor *executor, uint64_t seed, uint32_t domain, uint32_t path0,
                        uint32_t paths, uint32_t blocks, uint32_t *out, fohip_error *err);

/* The GBM step's binary64 exponential, out[i] = exp(x[i]) as the simulation kernels compute it
 * (a table-driven exp within one ulp of the correctly rounded value; parity probe). */
int fohip_gbm_exp_probe(fohip_executor *executor, const double *x, size_t n, double *out,
                        fohip_error *err);

/* ZIGNOR normals of paths path0..path0+paths-1: out[p * count + i]; draws[p * count + i] is the
 * number of u64 draws normal i consumed (1 on the fast path). */
int fohip_stream_normals(fohip_executor *executor, uint64_t seed, uint32_t domain, uint32_t path0,
                         uint32_t paths, uint32_t count, double *out, uint32_t *draws,
                         fohip_error *err);

/* The fit sweep's schedule for groups with the given fused-fit chunk counts (2048 training paths
 * each) and backward steps, teams of blocks holding at most `chunks_per_block` chunks each:
 * entries[8 e .. 8 e + 7] = (group, chunk_begin, chunk_end, members, team, leader, first partial
 * record low, high word), in the order the launch's blocks take them by ticket (teams by
 * decreasing cost, each team's members consecutively). Groups with no backward step get no
 * entry. Host logic only (no device needed), for tests and diagnostics; no counterpart in the
 * reference. FOHIP_E_INVALID_ARGUMENT when `entry_capacity` is too small (*entry_count then
 * holds the size needed). */
int fohip_sweep_schedule(const uint32_t *chunks, const uint32_t *fit_rows, uint32_t groups,
                         uint32_t chunks_per_block, uint32_t *entries, uint64_t entry_capacity,
                         uint64_t *entry_count, uint32_t *teams);

/* ---- host-side validation of the financial inputs (no device needed) ---- */
int fohip_validate_time_grid(const double *times, size_t nodes, fohip_error *err);
int fohip_validate_log_discount_curve(const double *times, const double *log_discounts,
                                      size_t nodes, fohip_error *err);
int fohip_validate_cumulative_yield_curve(const double *times, const double *cumulative_yields,
                                          size_t nodes, fohip_error *err);
int fohip_validate_event_schedule(const double *times, size_t nodes,
                                  const uint32_t *dividend_nodes, const double *dividend_amounts,
                                  size_t dividends, const uint32_t *exercise_nodes,
                                  size_t exercise_count, uint32_t expiry_node, fohip_error *err);
int fohip_validate_constant_volatility(double volatility, fohip_error *err);
int fohip_validate_cash_dividend_amount(double amount, fohip_error *err);
int fohip_validate_ragged_offsets(const int64_t *offsets, size_t offsets_len, size_t values_len,
                                  fohip_error *err);
/* Per-row checks of a ragged exercise schedule (times[offsets[r] - offsets[0] ..]): every time
 * finite and positive, then strictly increasing; the first failing row is reported, as
 * RaggedExerciseTimes in python/finoptions/batch_lsm.py does. The offsets must have passed
 * fohip_validate_ragged_offsets. */
int fohip_validate_ragged_exercise_times(const int64_t *offsets, size_t offsets_len,
                                         const double *times, fohip_error *err);
/* Per-row checks of ragged cash dividends (python/finoptions/dividends.py): times finite and
 * positive, then nondecreasing, then amounts finite, nonnegative and not -0.0. */
int fohip_validate_ragged_cash_dividends(const int64_t *offsets, size_t offsets_len,
                                         const double *times, const double *amounts,
                                         fohip_error *err);

#ifdef __cplusplus
}
#endif

#endif /* FOHIP_H */
