/*
 * CPU restatement of the reference Longstaff-Schwartz hot path.
 *
 * TEST INFRASTRUCTURE ONLY. This library is the parity oracle for the HIP engine: it is loaded by
 * tests/, by __graft_entry__.smoke() as the checker, and by bench.py's cpu_baseline leg. No product
 * code path may call it.
 *
 * Every function restates a reference function operation for operation (citations are to files of
 * jorgensandhaug/finoptions-engine). Floating-point semantics follow Rust on x86_64-linux-gnu:
 * binary64 SSE2 arithmetic with no contraction (`-ffp-contract=off`), `f64::mul_add` -> fma(),
 * `f64::exp`/`ln`/`hypot`/`sqrt` -> glibc exp()/log()/hypot()/sqrt().
 */
#ifndef LSM_ORACLE_H
#define LSM_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Error codes: the same numbering as include/fohip.h (FOHIP_E_*). */
typedef struct {
    int32_t code;
    int32_t phase;     /* 0 none, 1 training simulation, 2 fit, 3 evaluation simulation, 4 evaluate */
    int64_t row;       /* strike row (grouped) or -1 */
    int64_t node;      /* exercise node or -1 */
    int64_t path;      /* path or -1 */
    int64_t index;     /* auxiliary integer (interval, rank, observed count, ...) */
    double observed;   /* auxiliary value */
    double limit;      /* auxiliary value */
} or_error;

/* ---- random streams: philox.rs, normal.rs, normal_tables.rs ---- */
void or_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]);
void or_stream_u64(uint64_t seed, uint32_t domain, uint32_t path, uint64_t count, uint64_t *out);
/* Returns 0 on success; 1 budget exhausted; 2 rejection limit. Also reports the u64 draws
 * consumed per normal when `draws` is non-null. */
int or_stream_normals(uint64_t seed, uint32_t domain, uint32_t path, uint32_t count, double *out,
                      uint32_t *draws);
void or_symmetric_unit(const uint64_t *bits, size_t n, double *out);
void or_open_unit(const uint64_t *bits, size_t n, double *out);
void or_standard_unit(const uint64_t *bits, size_t n, double *out);
void or_zig_tables(double *x257, double *f257, double *r);

/* ---- GBM plan: path.rs:55-99, step.rs:27-95, dividend.rs:59-81 ---- */
typedef struct {
    uint32_t nodes;                /* grid length */
    const double *times;           /* [nodes] */
    const double *log_discounts;   /* [nodes] */
    const double *cum_yields;      /* [nodes] */
    double volatility;
    int32_t dividend_policy;       /* 0 liquidator, 1 survivor */
    uint32_t dividends;
    const uint32_t *dividend_nodes;
    const double *dividend_amounts;
    uint32_t exercise_count;
    const uint32_t *exercise_nodes;
    uint32_t expiry_node;
} or_plan_input;

int or_simulate_seeded(const or_plan_input *plan, double initial_spot, uint32_t paths,
                       uint64_t seed, uint32_t domain, double *out, or_error *err);
int or_simulate_normals(const or_plan_input *plan, double initial_spot, uint32_t paths,
                        const double *normals, double *out, or_error *err);
void or_step_coefficients(const or_plan_input *plan, double *log_drift, double *diffusion);

/* ---- regression.rs ---- */
typedef struct {
    int32_t degree;
    double mean, scale;
    double alpha[4], beta[4], coef[5], norm[5];
    uint32_t observations;
    uint32_t rank;
    double condition;
} or_model;

int or_fit_regression(const double *x, const double *y, uint32_t n, int32_t degree,
                      double rank_tolerance, double condition_limit, or_model *model,
                      or_error *err);
double or_predict(const or_model *model, double x);
void or_polynomial_moments(const double *z, const double *y, const double *b, size_t n,
                           double out[3]);
void or_recurrence_moments(const double *z, const double *y, const double *prev,
                           const double *cur, double *next, size_t n, double alpha, double beta,
                           double out[3]);

/* ---- policy.rs: fit / evaluate / price_seeded ---- */
typedef struct {
    double strike;
    int32_t degree;
    double rank_tolerance, condition_limit;
    int32_t tie_rule;              /* 0 strictly greater, 1 greater or equal */
    int32_t insufficient_support;  /* 0 fail, 1 continue */
    double tie_window;             /* near-tie count: |intrinsic - continuation| <= w |continuation| */
} or_estimator;

/* kind: 0 origin decision, 1 no ITM paths, 2 insufficient support, 3 regression */
typedef struct {
    uint32_t node;
    int32_t kind;
    int32_t exercise;
    double estimated_continuation;
    uint32_t itm_paths;
    uint32_t required_paths;
    or_model model;
    uint32_t near_ties;            /* rule applications at this node within the near-tie window */
} or_node_fit;

typedef struct {
    double price;
    uint32_t paths;
    int32_t has_sd;
    double sd, se;
    uint32_t obs;                  /* observation rows */
    uint32_t *observation_nodes;   /* [obs] caller buffer */
    uint32_t *exercise_counts;     /* [obs] caller buffer */
    or_node_fit *fits;             /* [exercise_count] caller buffer */
    uint32_t fit_count;
    uint64_t near_tie_training, near_tie_evaluation;
} or_call_result;

int or_price_lsm_call(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                      uint32_t training_paths, uint32_t evaluation_paths, uint64_t training_seed,
                      uint32_t training_domain, uint64_t evaluation_seed,
                      uint32_t evaluation_domain, or_call_result *result, or_error *err);

/* Fits on a caller-supplied time-major training matrix and evaluates on a caller-supplied
 * evaluation matrix (policy.rs:215-296, 642-736). Used with or_simulate_normals for fixtures. */
int or_fit_evaluate_paths(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                          const double *training, uint32_t training_paths,
                          const double *evaluation, uint32_t evaluation_paths,
                          or_call_result *result, or_error *err);

/* ---- batch/pricing.rs: group-addressed strikes ---- */
typedef struct {
    uint32_t groups;
    const uint32_t *group_offsets;     /* [groups + 1] */
    const double *initial_spots;       /* [groups] */
    const double *maturities;          /* [groups] */
    const double *rates;               /* [groups] */
    const double *volatilities;        /* [groups] */
    uint32_t rows;
    const double *strikes;             /* [rows] */
    int32_t dividend_model;            /* 0 continuous yield, 1 cash dividends, 2 escrowed cash */
    const double *dividend_yields;     /* [groups] */
    const int64_t *cash_offsets;       /* [groups + 1] */
    const double *cash_times;
    const double *cash_amounts;
    int32_t cash_policy;
    int32_t mesh_kind;                 /* 0 uniform, 1 cash-dividend dates, 2 explicit times */
    uint32_t base_intervals;
    int32_t origin_exercise;           /* 0 excluded, 1 included */
    const int64_t *exercise_offsets;   /* [groups + 1] (explicit) */
    const double *exercise_times;
    uint32_t training_paths, evaluation_paths;
    const uint32_t *training_path_counts;   /* [groups] or NULL (ragged extension) */
    const uint32_t *evaluation_path_counts; /* [groups] or NULL */
    uint64_t training_seed, evaluation_seed;
    uint32_t training_domain_start, evaluation_domain_start;
    int32_t degree;
    double rank_tolerance, condition_limit;
    int32_t tie_rule, insufficient_support;
    int32_t escrowed_volatility;       /* model 2: 0 prepaid forward, 1 spot matched at origin */
    double tie_window;                 /* near-tie window of the counts below */
} or_grouped_input;

typedef struct {
    double *prices, *sds, *ses;
    uint8_t *origin_fitted, *origin_exercised;
    uint32_t *regression_nodes, *no_support_nodes, *insufficient_nodes, *min_itm_paths;
    double *max_condition;
    uint32_t *early_paths, *expiry_paths;
    uint64_t *near_tie_training, *near_tie_evaluation;   /* optional (NULL skips) */
} or_batch_output;

/* Builds the transition grid of one group (pricing.rs:967-1114, 862-965). Returns the
 * node count, or -1 when `capacity` is too small. Exercise nodes go to exercise_nodes. */
int or_group_grid(const or_grouped_input *in, uint32_t group, double *times, uint32_t capacity,
                  uint32_t *exercise_nodes, uint32_t *exercise_count, uint32_t *div_nodes,
                  double *div_amounts, uint32_t *div_count);

/* threads of the single call's path-parallel simulation */
void or_set_threads(int threads);

/* threads <= 1: groups in order on the calling thread; otherwise group-parallel over `threads`
 * threads with every group serial (the reference's Rayon model); results do not depend on it. */
int or_price_grouped(const or_grouped_input *in, or_batch_output *out, int threads,
                     or_error *err);

/* Per-path frozen-policy pathwise sensitivities on the evaluation population of one group:
 * d price / d (S0, sigma, r, q) for every strike of the group, with the policy fitted on the
 * training population exactly as or_price_grouped does. */
int or_grouped_pathwise_greeks(const or_grouped_input *in, uint32_t group, double *delta,
                               double *vega, double *rho, double *dq, or_error *err);

/* The same sensitivities for one scheduled call (or_price_lsm_call's request); the rate and
 * yield directions shift the supplied curves in parallel. out[4] = delta, vega, rho, yield. */
int or_call_pathwise_greeks(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                            uint32_t training_paths, uint32_t evaluation_paths,
                            uint64_t training_seed, uint32_t training_domain,
                            uint64_t evaluation_seed, uint32_t evaluation_domain, double *out,
                            or_error *err);

/* Test utility: group g of eval_in priced under the policy fitted on group g of fit_in;
 * stop_mode 0 decide, 1 decide and record stops[strike * ne + path], 2 hold recorded rows. */
int or_grouped_frozen_prices(const or_grouped_input *fit_in, const or_grouped_input *eval_in, uint32_t g,
                             int32_t *stops, int stop_mode, double *prices, or_error *err);

/* The observation matrix [obs][paths] of group g's training (evaluation = 0) or evaluation
 * population, as the fit and the evaluation see it (escrowed groups: reconstructed stock). */
int or_grouped_paths(const or_grouped_input *in, uint32_t group, int evaluation, double *out,
                     or_error *err);

/* Escrowed prepaid cash per grid node of group g (path.rs:897-944): out[0] is the origin value,
 * out[node] the value observed at exercise nodes; returns the node count or a negative code. */
int or_escrowed_prepaid_cash(const or_grouped_input *in, uint32_t group, double *out,
                             uint32_t capacity, or_error *err);

/* Fitted models of one group (strikes x exercise rows) for coefficient-level parity checks. */
int or_grouped_models(const or_grouped_input *in, uint32_t group, or_node_fit *fits,
                      uint32_t capacity_per_strike, or_error *err);

#ifdef __cplusplus
}
#endif

#endif
