/*
 * CPU restatement of the reference Longstaff-Schwartz hot path (parity oracle).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/, by __graft_entry__.smoke() as the checker and by
 * bench.py's cpu_baseline leg. It is never the thing measured for the GPU, never shipped in the
 * product path, and never a fallback.
 *
 * Parity pins (see tests/test_oracle.py):
 *   Philox            Random123 known-answer vectors (philox.rs:121-130)
 *   uniforms          exact bit algebra vs. closed forms
 *   ZIGNOR normals    16-value bit pattern (normal.rs:116-141); table digest (tests/golden)
 *   moment kernels    exact values (runtime/src/lib.rs:482-539)
 *   dividends / step  exact cases (dividend.rs:98-121, step.rs:143-195)
 *   regression        numpy SVD least squares + condition (regression.rs:1181-1214),
 *                     degree-0 mean (regression.rs:1280-1293), affine rescaling (:1252-1278)
 *   fit / evaluate    European slice closed form (policy.rs:1355-1379), origin case = 100
 *                     exactly (policy.rs:1443-1465), support fixtures (policy.rs:1489-1566),
 *                     Black-Scholes with yield (tests/test_lsm.py:134-176), HHL one-dividend
 *                     reference (policy.rs:1617-1670)
 *   grouped batches   one-strike group == single call; duplicate strikes equal; grid builders
 *                     vs. closed forms (batch/pricing.rs tests)
 */
#define _GNU_SOURCE
#include "lsm_oracle.h"

#include <math.h>
#include <float.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* error codes (identical numbering to include/fohip.h) */
enum {
    E_OK = 0,
    E_NON_FINITE_TRANSITION = 200,
    E_NORMAL_REJECTION_LIMIT = 201,
    E_PATH_NAMESPACE_OVERFLOW = 202,
    E_EMPTY_PATH_SET = 203,
    E_INSUFFICIENT_OBSERVATIONS = 300,
    E_DEGENERATE_PREDICTOR = 301,
    E_RANK_DEFICIENT = 302,
    E_CONDITION_LIMIT = 303,
    E_NON_FINITE_FIT = 304,
    E_JACOBI_NOT_CONVERGED = 305,
    E_NON_FINITE_PREDICTOR = 306,
    E_NON_FINITE_RESPONSE = 307,
    E_NON_FINITE_PREDICTION = 308,
    E_RANDOM_DOMAIN_COLLISION = 400,
    E_NON_FINITE_DISCOUNTED_PAYOFF = 401,
    E_OUTSIDE_TRAINING_SUPPORT = 402,
    E_UNREPRESENTABLE_DISCOUNT = 403,
    E_PATH_COUNT_OUTSIDE_DOMAIN = 404,
    E_ESCROWED_PREPAID_FORWARD = 508,
    E_INTERNAL = 904,
};

enum { PH_NONE = 0, PH_TRAIN_SIM = 1, PH_FIT = 2, PH_EVAL_SIM = 3, PH_EVALUATE = 4 };

static int set_err(or_error *e, int code, int phase, int64_t node, int64_t path, int64_t index,
                   double observed, double limit) {
    if (e) {
        e->code = code;
        e->phase = phase;
        e->row = -1;
        e->node = node;
        e->path = path;
        e->index = index;
        e->observed = observed;
        e->limit = limit;
    }
    return code;
}

/* =============================== Philox4x32-10 (philox.rs:3-48) =============================== */

#define PHILOX_M0 0xD2511F53u
#define PHILOX_M1 0xCD9E8D57u
#define PHILOX_W0 0x9E3779B9u
#define PHILOX_W1 0xBB67AE85u

/* philox.rs:18-29: one round; output lanes [hi1^c1^k0, lo1, hi0^c3^k1, lo0]. */
static inline void philox_round(uint32_t c[4], const uint32_t k[2]) {
    uint64_t p0 = (uint64_t)PHILOX_M0 * c[0];
    uint64_t p1 = (uint64_t)PHILOX_M1 * c[2];
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c[1] ^ k[0];
    uint32_t n2 = hi0 ^ c[3] ^ k[1];
    c[0] = n0;
    c[1] = lo1;
    c[2] = n2;
    c[3] = lo0;
}

/* philox.rs:32-41: ten rounds, key bumped before rounds 2..10. */
void or_philox4x32_10(const uint32_t counter[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {counter[0], counter[1], counter[2], counter[3]};
    uint32_t k[2] = {key[0], key[1]};
    philox_round(c, k);
    for (int i = 1; i < 10; ++i) {
        k[0] += PHILOX_W0;
        k[1] += PHILOX_W1;
        philox_round(c, k);
    }
    memcpy(out, c, sizeof c);
}

/* philox.rs:52-114: one path-local stream; counter [blk_lo, blk_hi, path, domain]. */
typedef struct {
    uint32_t key[2];
    uint32_t domain, path;
    uint64_t next_block;
    uint32_t remaining;
    uint32_t buffer[4];
    int position;
    uint64_t draws;
} stream_t;

static void stream_init(stream_t *s, uint64_t seed, uint32_t domain, uint32_t path,
                        uint32_t budget) {
    s->key[0] = (uint32_t)seed;
    s->key[1] = (uint32_t)(seed >> 32);
    s->domain = domain;
    s->path = path;
    s->next_block = 0;
    s->remaining = budget;
    s->position = 4;
    s->draws = 0;
}

static inline uint32_t stream_next_u32(stream_t *s) {
    if (s->position == 4) {
        uint32_t ctr[4] = {(uint32_t)s->next_block, (uint32_t)(s->next_block >> 32), s->path,
                           s->domain};
        or_philox4x32_10(ctr, s->key, s->buffer);
        s->position = 0;
        s->next_block += 1;
    }
    return s->buffer[s->position++];
}

static inline uint64_t stream_next_u64(stream_t *s) {
    uint64_t lo = stream_next_u32(s);
    uint64_t hi = stream_next_u32(s);
    s->draws += 1;
    return lo | (hi << 32);
}

void or_stream_u64(uint64_t seed, uint32_t domain, uint32_t path, uint64_t count, uint64_t *out) {
    stream_t s;
    stream_init(&s, seed, domain, path, 0);
    for (uint64_t i = 0; i < count; ++i) out[i] = stream_next_u64(&s);
}

/* ======================== ZIGNOR normals (normal.rs, normal_tables.rs) ======================== */

static double ZX[257], ZF[257], ZR;
static int zig_ready = 0;

static double decimal18(double v) {
    /* The published tables are the 18-decimal renderings of the construction below. */
    char buf[64];
    snprintf(buf, sizeof buf, "%.18f", v);
    return strtod(buf, NULL);
}

/* Doornik (2005) ZIGNOR construction with 256 layers, r = 3.6541528853610088 and layer area
 * v = 0.00492867323399: x0 = v / f(r), x1 = r, x_i = f^-1(v / x_{i-1} + f(x_{i-1})), x256 = 0,
 * F_i = f(x_i), f(x) = exp(-x^2 / 2). One published entry (F[94]) was rendered on a platform whose
 * exp() differs by one unit in the 18th decimal; it is restated as published. */
static void zig_build(void) {
    if (zig_ready) return;
    const double r = 3.6541528853610088, v = 0.00492867323399;
    double x[257];
    x[0] = v / exp(-r * r / 2.0);
    x[1] = r;
    for (int i = 2; i < 256; ++i) {
        double last = x[i - 1];
        x[i] = sqrt(-2.0 * log(v / last + exp(-last * last / 2.0)));
    }
    x[256] = 0.0;
    for (int i = 0; i < 257; ++i) {
        ZX[i] = decimal18(x[i]);
        ZF[i] = decimal18(exp(-x[i] * x[i] / 2.0));
    }
    ZF[94] = strtod("0.206405406398679298", NULL);
    ZR = decimal18(r);
    zig_ready = 1;
}

void or_zig_tables(double *x257, double *f257, double *r) {
    zig_build();
    memcpy(x257, ZX, sizeof ZX);
    memcpy(f257, ZF, sizeof ZF);
    *r = ZR;
}

#define FRACTION_MASK ((UINT64_C(1) << 52) - 1)
#define EXPONENT_0 (UINT64_C(1023) << 52)
#define EXPONENT_1 (UINT64_C(1024) << 52)
#define MAX_REJECTION_ITERATIONS 1048576

static inline double from_bits(uint64_t b) {
    double d;
    memcpy(&d, &b, sizeof d);
    return d;
}

/* normal.rs:33-35 */
static inline double symmetric_unit(uint64_t bits) {
    return from_bits(((bits >> 12) & FRACTION_MASK) | EXPONENT_1) - 3.0;
}
/* normal.rs:38-43 */
static inline double open_unit_bits(uint64_t bits) {
    return from_bits(((bits >> 12) & FRACTION_MASK) | EXPONENT_0) - (1.0 - DBL_EPSILON / 2.0);
}
/* normal.rs:46-48 */
static inline double standard_unit_bits(uint64_t bits) {
    return (double)(bits >> 11) * (1.0 / 9007199254740992.0);
}

void or_symmetric_unit(const uint64_t *bits, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = symmetric_unit(bits[i]);
}
void or_open_unit(const uint64_t *bits, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = open_unit_bits(bits[i]);
}
void or_standard_unit(const uint64_t *bits, size_t n, double *out) {
    for (size_t i = 0; i < n; ++i) out[i] = standard_unit_bits(bits[i]);
}

/* normal.rs:50-63 */
static int sample_tail(stream_t *s, double signed_unit, double *out) {
    for (int it = 0; it < MAX_REJECTION_ITERATIONS; ++it) {
        double x = log(open_unit_bits(stream_next_u64(s))) / ZR;
        double y = log(open_unit_bits(stream_next_u64(s)));
        if (-2.0 * y >= x * x) {
            *out = signed_unit < 0.0 ? x - ZR : ZR - x;
            return 0;
        }
    }
    return 2;
}

/* normal.rs:66-91 */
static int sample_rejected(stream_t *s, int index, double unit, double value, double *out) {
    for (int it = 0; it < MAX_REJECTION_ITERATIONS; ++it) {
        if (index == 0) return sample_tail(s, unit, out);
        double ordinate = ZF[index + 1] + (ZF[index] - ZF[index + 1]) *
                                              standard_unit_bits(stream_next_u64(s));
        if (ordinate < exp(-0.5 * value * value)) {
            *out = value;
            return 0;
        }
        uint64_t bits = stream_next_u64(s);
        index = (int)(bits & 0xff);
        unit = symmetric_unit(bits);
        value = unit * ZX[index];
        if (fabs(value) < ZX[index + 1]) {
            *out = value;
            return 0;
        }
    }
    return 2;
}

/* normal.rs:96-109. Returns 0, 1 (budget exhausted) or 2 (rejection limit). */
static inline int sample_normal(stream_t *s, double *out) {
    if (s->remaining == 0) return 1;
    s->remaining -= 1;
    uint64_t bits = stream_next_u64(s);
    int index = (int)(bits & 0xff);
    double unit = symmetric_unit(bits);
    double value = unit * ZX[index];
    if (fabs(value) < ZX[index + 1]) {
        *out = value;
        return 0;
    }
    return sample_rejected(s, index, unit, value, out);
}

int or_stream_normals(uint64_t seed, uint32_t domain, uint32_t path, uint32_t count, double *out,
                      uint32_t *draws) {
    zig_build();
    stream_t s;
    stream_init(&s, seed, domain, path, count);
    for (uint32_t i = 0; i < count; ++i) {
        uint64_t before = s.draws;
        int rc = sample_normal(&s, &out[i]);
        if (rc) return rc;
        if (draws) draws[i] = (uint32_t)(s.draws - before);
    }
    return 0;
}

/* ============================ GBM plan (path.rs:55-99, step.rs) ============================ */

typedef struct {
    uint32_t nodes, intervals, obs;
    double *log_drift, *diffusion;   /* [intervals] */
    double *dividend;                /* [nodes], NaN when absent */
    uint32_t *obs_nodes;             /* [obs] */
    int policy;
} plan_t;

/* step.rs:27-57: var = fma(s, s, 0) * dt; drift = -dlogD - dQ - 0.5 var; diff = sqrt(var). */
static void plan_build(const or_plan_input *in, plan_t *p) {
    p->nodes = in->nodes;
    p->intervals = in->nodes - 1;
    p->obs = in->exercise_count + 1;
    p->log_drift = malloc(sizeof(double) * p->intervals);
    p->diffusion = malloc(sizeof(double) * p->intervals);
    p->dividend = malloc(sizeof(double) * in->nodes);
    p->obs_nodes = malloc(sizeof(uint32_t) * p->obs);
    p->policy = in->dividend_policy;
    const double sigma = in->volatility;
    for (uint32_t to = 1; to < in->nodes; ++to) {
        uint32_t from = to - 1;
        double dt = in->times[to] - in->times[from];
        double interval_log_discount = in->log_discounts[to] - in->log_discounts[from];
        double interval_yield = in->cum_yields[to] - in->cum_yields[from];
        double variance = fma(sigma, sigma, 0.0) * dt;
        double log_drift = -interval_log_discount - interval_yield - 0.5 * variance;
        p->log_drift[from] = log_drift;
        p->diffusion[from] = sqrt(variance);
    }
    for (uint32_t i = 0; i < in->nodes; ++i) p->dividend[i] = NAN;
    for (uint32_t d = 0; d < in->dividends; ++d)
        p->dividend[in->dividend_nodes[d]] = in->dividend_amounts[d];
    for (uint32_t e = 0; e < in->exercise_count; ++e) p->obs_nodes[e] = in->exercise_nodes[e];
    p->obs_nodes[in->exercise_count] = in->expiry_node;
}

static void plan_free(plan_t *p) {
    free(p->log_drift);
    free(p->diffusion);
    free(p->dividend);
    free(p->obs_nodes);
}

void or_step_coefficients(const or_plan_input *plan, double *log_drift, double *diffusion) {
    plan_t p;
    plan_build(plan, &p);
    memcpy(log_drift, p.log_drift, sizeof(double) * p.intervals);
    memcpy(diffusion, p.diffusion, sizeof(double) * p.intervals);
    plan_free(&p);
}

/* dividend.rs:59-81 (inputs here are always finite and nonnegative). */
static inline double apply_dividend(int policy, double amount, double spot) {
    if (policy == 0) {
        double after = spot - amount;
        return after > 0.0 ? after : 0.0;   /* f64::max(after, 0.0) for finite after */
    }
    return spot > amount ? spot - amount : spot;
}

/* step.rs:89-95 */
static inline double advance(const plan_t *p, uint32_t interval, double spot, double normal) {
    if (spot == 0.0) return 0.0;
    double exponent = fma(p->diffusion[interval], normal, p->log_drift[interval]);
    return spot * exp(exponent);
}

/* seeded.rs:399-459 with an external normal source when `normals` is non-null (path.rs:384-531
 * `simulate` applies the identical event ordering to a supplied [intervals, paths] matrix). */
static int simulate_one(const plan_t *p, double initial_spot, uint32_t path, uint64_t seed,
                        uint32_t domain, const double *normals, uint32_t paths, double *obs_out,
                        size_t obs_stride, or_error *err, int phase) {
    stream_t s;
    if (!normals) stream_init(&s, seed, domain, path, p->intervals);
    double spot = initial_spot;
    uint32_t observation = 0;
    if (p->obs_nodes[0] == 0) {
        obs_out[0] = spot;
        observation = 1;
    }
    const uint32_t terminal = p->obs_nodes[p->obs - 1];
    for (uint32_t interval = 0; interval < p->intervals; ++interval) {
        double normal;
        if (normals) {
            normal = normals[(size_t)interval * paths + path];
        } else {
            int rc = sample_normal(&s, &normal);
            if (rc == 2)
                return set_err(err, E_NORMAL_REJECTION_LIMIT, phase, -1, path, interval, 0, 0);
            if (rc == 1) return set_err(err, E_INTERNAL, phase, -1, path, interval, 0, 0);
        }
        double before = spot;
        double next = advance(p, interval, before, normal);
        if (!isfinite(next))
            return set_err(err, E_NON_FINITE_TRANSITION, phase, interval + 1, path, interval,
                           before, normal);
        spot = next;
        uint32_t node = interval + 1;
        double cash = p->dividend[node];
        int has_cash = !isnan(cash);
        int terminal_cash = has_cash && node == terminal;
        if (terminal_cash) spot = apply_dividend(p->policy, cash, spot);
        if (observation < p->obs && p->obs_nodes[observation] == node) {
            obs_out[(size_t)observation * obs_stride] = spot;
            observation += 1;
        }
        if (has_cash && !terminal_cash) spot = apply_dividend(p->policy, cash, spot);
    }
    return 0;
}

static int simulate_population(const plan_t *p, double initial_spot, uint32_t paths, uint64_t seed,
                               uint32_t domain, const double *normals, double *out, or_error *err,
                               int phase) {
    zig_build();
    if (paths == 0) return set_err(err, E_EMPTY_PATH_SET, phase, -1, -1, 0, 0, 0);
    int64_t failed = -1;
    or_error first = {0};
#pragma omp parallel for schedule(static)
    for (int64_t path = 0; path < (int64_t)paths; ++path) {
        or_error local = {0};
        int rc = simulate_one(p, initial_spot, (uint32_t)path, seed, domain, normals, paths,
                              out + path, paths, &local, phase);
        if (rc) {
#pragma omp critical
            {
                if (failed < 0 || path < failed) {
                    failed = path;
                    first = local;
                }
            }
        }
    }
    if (failed >= 0) {
        if (err) *err = first;
        return first.code;
    }
    return 0;
}

int or_simulate_seeded(const or_plan_input *plan, double initial_spot, uint32_t paths,
                       uint64_t seed, uint32_t domain, double *out, or_error *err) {
    plan_t p;
    plan_build(plan, &p);
    int rc = simulate_population(&p, initial_spot, paths, seed, domain, NULL, out, err, PH_NONE);
    plan_free(&p);
    return rc;
}

int or_simulate_normals(const or_plan_input *plan, double initial_spot, uint32_t paths,
                        const double *normals, double *out, or_error *err) {
    plan_t p;
    plan_build(plan, &p);
    int rc = simulate_population(&p, initial_spot, paths, 0, 0, normals, out, err, PH_NONE);
    plan_free(&p);
    return rc;
}

/* ===================================== regression.rs ===================================== */

#define JACOBI_MAX_SWEEPS 32
#define JACOBI_TOL (16.0 * DBL_EPSILON)

/* runtime/src/lib.rs:84-95 */
void or_polynomial_moments(const double *z, const double *y, const double *b, size_t n,
                           double out[3]) {
    double norm = 0.0, projection = 0.0, weighted = 0.0;
    for (size_t i = 0; i < n; ++i) {
        double sq = b[i] * b[i];
        norm += sq;
        projection = fma(y[i], b[i], projection);
        weighted = fma(z[i], sq, weighted);
    }
    out[0] = norm;
    out[1] = projection;
    out[2] = weighted;
}

/* runtime/src/lib.rs:141-158 */
void or_recurrence_moments(const double *z, const double *y, const double *prev,
                           const double *cur, double *next, size_t n, double alpha, double beta,
                           double out[3]) {
    double norm = 0.0, projection = 0.0, weighted = 0.0;
    for (size_t i = 0; i < n; ++i) {
        next[i] = fma(z[i] - alpha, cur[i], -beta * prev[i]);
        double sq = next[i] * next[i];
        norm += sq;
        projection = fma(y[i], next[i], projection);
        weighted = fma(z[i], sq, weighted);
    }
    out[0] = norm;
    out[1] = projection;
    out[2] = weighted;
}

/* regression.rs:711-750 */
static int orthonormal_transform(int degree, const double *alphas, const double *betas,
                                 const double *norms, double t[25]) {
    double previous[5] = {0}, current[5] = {0};
    current[0] = 1.0;
    memset(t, 0, sizeof(double) * 25);
    for (int k = 0; k <= degree; ++k) {
        double inverse_norm = 1.0 / sqrt(norms[k]);
        if (!isfinite(inverse_norm)) return k + 1;
        for (int m = 0; m <= k; ++m) t[k * 5 + m] = current[m] * inverse_norm;
        if (k == degree) break;
        double next[5] = {0};
        for (int m = 0; m <= k; ++m) {
            next[m + 1] += current[m];
            next[m] -= alphas[k] * current[m];
            next[m] -= betas[k] * previous[m];
        }
        for (int m = 0; m < 5; ++m)
            if (!isfinite(next[m])) return k + 2;
        memcpy(previous, current, sizeof current);
        memcpy(current, next, sizeof next);
    }
    return 0;
}

static int cmp_desc(const void *a, const void *b) {
    double x = *(const double *)a, y = *(const double *)b;
    return x < y ? 1 : (x > y ? -1 : 0);
}

/* regression.rs:766-832 (returns 0, or 1 + basis degree for NonFiniteFit, or -1 when the sweep
 * limit is reached). */
static int jacobi_singular_values(double m[25], int columns, double sv[5]) {
    int converged = 0;
    for (int sweep = 0; sweep < JACOBI_MAX_SWEEPS; ++sweep) {
        converged = 1;
        for (int left = 0; left < columns - 1; ++left) {
            for (int right = left + 1; right < columns; ++right) {
                double ln = 0.0, rn = 0.0, cross = 0.0;
                for (int row = 0; row < columns; ++row) {
                    double lv = m[left * 5 + row], rv = m[right * 5 + row];
                    ln = fma(lv, lv, ln);
                    rn = fma(rv, rv, rn);
                    cross = fma(lv, rv, cross);
                }
                if (!isfinite(ln) || !isfinite(rn) || !isfinite(cross) || ln <= 0.0 || rn <= 0.0)
                    return 1 + right;
                if (fabs(cross) <= JACOBI_TOL * sqrt(ln * rn)) continue;
                converged = 0;
                double zeta = (rn - ln) / (2.0 * cross);
                double tangent;
                if (zeta == 0.0) {
                    tangent = 1.0;
                } else {
                    double sign = signbit(zeta) ? -1.0 : 1.0;
                    tangent = sign / (fabs(zeta) + hypot(zeta, 1.0));
                }
                double cosine = 1.0 / sqrt(fma(tangent, tangent, 1.0));
                double sine = cosine * tangent;
                for (int row = 0; row < columns; ++row) {
                    double lv = m[left * 5 + row], rv = m[right * 5 + row];
                    m[left * 5 + row] = fma(cosine, lv, -sine * rv);
                    m[right * 5 + row] = fma(sine, lv, cosine * rv);
                }
            }
        }
        if (converged) break;
    }
    if (!converged) return -1;
    for (int c = 0; c < 5; ++c) sv[c] = 0.0;
    for (int c = 0; c < columns; ++c) {
        double sq = 0.0;
        for (int row = 0; row < columns; ++row) {
            double v = m[c * 5 + row];
            sq += v * v;
        }
        sv[c] = sqrt(sq);
    }
    qsort(sv, (size_t)columns, sizeof(double), cmp_desc);
    return 0;
}

/* regression.rs:432-523 (fit_orthogonal) + 540-577 (diagnose_design) on a standardized column. */
static int fit_standardized(const double *z, const double *y, uint32_t n, int degree,
                            double rank_tol, double cond_limit, or_model *m, or_error *err) {
    double *prev = calloc(n, sizeof(double));
    double *cur = malloc(sizeof(double) * n);
    double *next = malloc(sizeof(double) * n);
    for (uint32_t i = 0; i < n; ++i) cur[i] = 1.0;
    int rc = 0;
    double previous_norm = 1.0;
    double mom[3];
    for (int k = 0; k < 4; ++k) m->alpha[k] = m->beta[k] = 0.0;
    for (int k = 0; k < 5; ++k) m->coef[k] = m->norm[k] = 0.0;
    or_polynomial_moments(z, y, cur, n, mom);
    for (int k = 0; k <= degree; ++k) {
        double norm = mom[0], projection = mom[1], weighted = mom[2];
        if (!isfinite(norm) || !isfinite(projection) || !isfinite(weighted)) {
            rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, k, 0, 0);
            goto done;
        }
        if (norm <= 0.0 || sqrt(norm / (double)n) <= rank_tol) {
            rc = set_err(err, E_RANK_DEFICIENT, PH_FIT, -1, -1, k, (double)(degree + 1), rank_tol);
            goto done;
        }
        m->norm[k] = norm;
        m->coef[k] = projection / norm;
        if (!isfinite(m->coef[k])) {
            rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, k, 0, 0);
            goto done;
        }
        if (k < degree) {
            m->alpha[k] = weighted / norm;
            m->beta[k] = k == 0 ? 0.0 : norm / previous_norm;
            if (!isfinite(m->alpha[k]) || !isfinite(m->beta[k])) {
                rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, k, 0, 0);
                goto done;
            }
            or_recurrence_moments(z, y, prev, cur, next, n, m->alpha[k], m->beta[k], mom);
            if (!isfinite(mom[0]) || !isfinite(mom[1]) || !isfinite(mom[2])) {
                rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, k + 1, 0, 0);
                goto done;
            }
            previous_norm = norm;
            double *t = prev;
            prev = cur;
            cur = next;
            next = t;
        }
    }
    {
        double t[25], sv[5];
        int columns = degree + 1;
        int bad = orthonormal_transform(degree, m->alpha, m->beta, m->norm, t);
        if (bad) {
            rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, bad - 1, 0, 0);
            goto done;
        }
        int j = jacobi_singular_values(t, columns, sv);
        if (j < 0) {
            rc = set_err(err, E_JACOBI_NOT_CONVERGED, PH_FIT, -1, -1, columns, 0, 0);
            goto done;
        }
        if (j > 0) {
            rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, j - 1, 0, 0);
            goto done;
        }
        double largest = sv[0], threshold = rank_tol * largest;
        int rank = 0;
        for (int c = 0; c < columns; ++c)
            if (sv[c] > threshold) rank += 1;
        if (rank < columns) {
            rc = set_err(err, E_RANK_DEFICIENT, PH_FIT, -1, -1, rank, (double)columns, rank_tol);
            goto done;
        }
        double condition = largest / sv[columns - 1];
        if (!isfinite(condition)) {
            rc = set_err(err, E_NON_FINITE_FIT, PH_FIT, -1, -1, degree, 0, 0);
            goto done;
        }
        if (condition > cond_limit) {
            rc = set_err(err, E_CONDITION_LIMIT, PH_FIT, -1, -1, degree, condition, cond_limit);
            goto done;
        }
        m->rank = (uint32_t)rank;
        m->condition = condition;
    }
done:
    free(prev);
    free(cur);
    free(next);
    return rc;
}

/* Welford state of RegressionPopulation::push_indexed (regression.rs:279-297). */
typedef struct {
    double count, mean, m2;
} welford_t;

static inline void welford_push(welford_t *w, int degree, double x) {
    if (degree == 0) return;
    w->count += 1.0;
    double delta = x - w->mean;
    w->mean += delta / w->count;
    w->m2 = fma(delta, x - w->mean, w->m2);
}

/* fit_population (regression.rs:369-386) with the population's running location/scale. */
static int fit_population(const double *x, const double *y, uint32_t n, const welford_t *w,
                          int degree, double rank_tol, double cond_limit, or_model *m,
                          or_error *err) {
    if (n < (uint32_t)(degree + 1))
        return set_err(err, E_INSUFFICIENT_OBSERVATIONS, PH_FIT, -1, -1, n, degree + 1, degree);
    for (uint32_t i = 0; i < n; ++i)
        if (!isfinite(x[i])) return set_err(err, E_NON_FINITE_PREDICTOR, PH_FIT, -1, i, i, x[i], 0);
    for (uint32_t i = 0; i < n; ++i)
        if (!isfinite(y[i])) return set_err(err, E_NON_FINITE_RESPONSE, PH_FIT, -1, i, i, y[i], 0);
    double mean = 0.0, scale = 1.0;
    if (degree != 0) {
        scale = sqrt(w->m2 / w->count);
        if (!isfinite(w->mean) || !isfinite(scale) || scale <= 0.0)
            return set_err(err, E_DEGENERATE_PREDICTOR, PH_FIT, -1, -1, degree, scale, 0);
        mean = w->mean;
    }
    double *z = malloc(sizeof(double) * n);
    for (uint32_t i = 0; i < n; ++i) z[i] = degree == 0 ? 0.0 : (x[i] - mean) / scale;
    m->degree = degree;
    m->mean = mean;
    m->scale = scale;
    m->observations = n;
    int rc = fit_standardized(z, y, n, degree, rank_tol, cond_limit, m, err);
    free(z);
    return rc;
}

int or_fit_regression(const double *x, const double *y, uint32_t n, int32_t degree,
                      double rank_tolerance, double condition_limit, or_model *model,
                      or_error *err) {
    welford_t w = {0, 0, 0};
    for (uint32_t i = 0; i < n; ++i)
        if (isfinite(x[i])) welford_push(&w, degree, x[i]);
    return fit_population(x, y, n, &w, degree, rank_tolerance, condition_limit, model, err);
}

/* regression.rs:197-224 */
double or_predict(const or_model *m, double predictor) {
    double z = m->degree == 0 ? 0.0 : (predictor - m->mean) / m->scale;
    double prediction = m->coef[0], previous = 0.0, current = 1.0;
    for (int k = 0; k < m->degree; ++k) {
        double next = fma(z - m->alpha[k], current, -m->beta[k] * previous);
        prediction = fma(m->coef[k + 1], next, prediction);
        previous = current;
        current = next;
    }
    return prediction;
}

/* ======================================= policy.rs ======================================= */

enum { RULE_ORIGIN = 0, RULE_NO_SUPPORT = 1, RULE_CONTINUE = 2, RULE_REGRESSION = 3 };

static inline int should_exercise(int tie, double intrinsic, double continuation) {
    return tie == 0 ? intrinsic > continuation : intrinsic >= continuation;
}

/* Near tie of a rule application: the decisions a last-bit change of the continuation could
 * flip. Counting them is this project's addition; the reference has no such diagnostic. */
static inline int near_tie(double intrinsic, double continuation, double window) {
    return fabs(intrinsic - continuation) <= window * fabs(continuation);
}

static inline double call_intrinsic(double spot, double strike) {
    double v = spot - strike;
    return v > 0.0 ? v : 0.0;   /* f64::max(v, 0.0) for non-NaN v */
}

/* Fits one strike on the time-major training matrix (policy.rs:215-296). `fits` receives one
 * entry per exercise row in row order; `log_discounts` is the curve on the model grid. */
static int fit_policy(const plan_t *p, const double *log_discounts, double initial_spot,
                      const or_estimator *est, const double *training, uint32_t n,
                      or_node_fit *fits, double *values, double *xbuf, double *ybuf,
                      uint32_t *ibuf, or_error *err) {
    const uint32_t rows = p->obs - 1;
    const double strike = est->strike;
    const double *terminal = training + (size_t)rows * n;
    for (uint32_t i = 0; i < n; ++i) values[i] = call_intrinsic(terminal[i], strike);
    uint32_t later = p->obs_nodes[rows];
    for (int64_t r = (int64_t)rows - 1; r >= 0; --r) {
        uint32_t node = p->obs_nodes[r];
        or_node_fit *f = &fits[r];
        memset(f, 0, sizeof *f);
        f->node = node;
        if (node == 0) continue;
        /* discount_values (policy.rs:1049-1063) */
        double ld = log_discounts[later] - log_discounts[node];
        double discount = exp(ld);
        if (!isfinite(discount) || discount <= 0.0)
            return set_err(err, E_UNREPRESENTABLE_DISCOUNT, PH_FIT, node, -1, later, ld, 0);
        for (uint32_t i = 0; i < n; ++i) {
            values[i] *= discount;
            if (!isfinite(values[i]))
                return set_err(err, E_NON_FINITE_DISCOUNTED_PAYOFF, PH_FIT, node, i, 0, 0, 0);
        }
        const double *spots = training + (size_t)r * n;
        /* fit_positive_time_rule (policy.rs:471-540) */
        welford_t w = {0, 0, 0};
        uint32_t count = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (spots[i] > strike) {
                double x = spots[i] / strike;
                xbuf[count] = x;
                ybuf[count] = values[i];
                ibuf[count] = i;
                count += 1;
                if (isfinite(x)) welford_push(&w, est->degree, x);
            }
        }
        uint32_t required = (uint32_t)est->degree + 1;
        if (count == 0) {
            if (est->insufficient_support == 1) {
                f->kind = 2;
                f->itm_paths = 0;
                f->required_paths = required;
            } else {
                f->kind = 1;
            }
        } else if (count < required && est->insufficient_support == 1) {
            f->kind = 2;
            f->itm_paths = count;
            f->required_paths = required;
        } else {
            int rc = fit_population(xbuf, ybuf, count, &w, est->degree, est->rank_tolerance,
                                    est->condition_limit, &f->model, err);
            if (rc) {
                err->phase = PH_FIT;
                err->node = node;
                return rc;
            }
            f->kind = 3;
            f->itm_paths = count;
            /* apply_training_rule_for (policy.rs:570-601) */
            for (uint32_t j = 0; j < count; ++j) {
                uint32_t path = ibuf[j];
                double intrinsic = spots[path] - strike;
                double continuation = or_predict(&f->model, xbuf[j]);
                if (!isfinite(continuation))
                    return set_err(err, E_NON_FINITE_PREDICTION, PH_FIT, node, path, 0, xbuf[j],
                                   0);
                f->near_ties += (uint32_t)near_tie(intrinsic, continuation, est->tie_window);
                if (should_exercise(est->tie_rule, intrinsic, continuation))
                    values[path] = intrinsic;
            }
        }
        later = node;
    }
    /* discount_values_to_origin (policy.rs:1065-1079) */
    {
        double ld = log_discounts[later];
        double discount = exp(ld);
        if (!isfinite(discount) || discount <= 0.0)
            return set_err(err, E_UNREPRESENTABLE_DISCOUNT, PH_FIT, 0, -1, later, ld, 0);
        for (uint32_t i = 0; i < n; ++i) {
            values[i] *= discount;
            if (!isfinite(values[i]))
                return set_err(err, E_NON_FINITE_DISCOUNTED_PAYOFF, PH_FIT, 0, i, 0, 0, 0);
        }
    }
    if (p->obs_nodes[0] == 0) {
        /* online_mean (policy.rs:1110-1122) */
        double mean = 0.0;
        uint32_t count = 0;
        for (uint32_t i = 0; i < n; ++i) {
            count += 1;
            mean += (values[i] - mean) / (double)count;
        }
        double intrinsic = call_intrinsic(initial_spot, est->strike);
        fits[0].kind = 0;
        fits[0].node = 0;
        fits[0].exercise = should_exercise(est->tie_rule, intrinsic, mean);
        fits[0].estimated_continuation = mean;
    }
    return 0;
}

typedef struct {
    double price, sd, se;
    int has_sd;
    uint32_t *counts;   /* [obs] */
    uint64_t near_ties;
} eval_t;

/* FittedLsmCallPolicy::evaluate (policy.rs:642-736) + from_path_payoffs (:891-926). */
static int evaluate_policy(const plan_t *p, const double *log_discounts, double initial_spot,
                           const or_estimator *est, const or_node_fit *fits,
                           const double *evaluation, uint32_t n, double *payoffs, uint8_t *active,
                           eval_t *out, or_error *err) {
    const uint32_t rows = p->obs - 1;
    for (uint32_t r = 0; r <= rows; ++r) out->counts[r] = 0;
    if (rows > 0 && fits[0].kind == 0 && fits[0].exercise) {
        out->counts[0] = n;
        out->price = call_intrinsic(initial_spot, est->strike);
        out->sd = out->se = 0.0;
        out->has_sd = 1;
        return 0;
    }
    for (uint32_t i = 0; i < n; ++i) {
        active[i] = 1;
        payoffs[i] = 0.0;
    }
    for (uint32_t r = 0; r < rows; ++r) {
        const or_node_fit *f = &fits[r];
        const double *spots = evaluation + (size_t)r * n;
        uint32_t node = p->obs_nodes[r];
        if (f->kind == 0 || f->kind == 2) continue;
        if (f->kind == 1) {
            uint32_t itm = 0;
            for (uint32_t i = 0; i < n; ++i)
                if (active[i] && spots[i] > est->strike) itm += 1;
            if (itm > 0)
                return set_err(err, E_OUTSIDE_TRAINING_SUPPORT, PH_EVALUATE, node, -1, itm, 0, 0);
            continue;
        }
        double discount = exp(log_discounts[node]);
        uint32_t count = 0;
        for (uint32_t i = 0; i < n; ++i) {
            double spot = spots[i];
            if (!active[i] || spot <= est->strike) continue;
            double intrinsic = spot - est->strike;
            double continuation = or_predict(&f->model, spot / est->strike);
            if (!isfinite(continuation))
                return set_err(err, E_NON_FINITE_PREDICTION, PH_EVALUATE, node, i, 0,
                               spot / est->strike, 0);
            out->near_ties += (uint64_t)near_tie(intrinsic, continuation, est->tie_window);
            if (should_exercise(est->tie_rule, intrinsic, continuation)) {
                double payoff = intrinsic * discount;
                if (!isfinite(payoff))
                    return set_err(err, E_NON_FINITE_DISCOUNTED_PAYOFF, PH_EVALUATE, node, i, 0, 0,
                                   0);
                payoffs[i] = payoff;
                active[i] = 0;
                count += 1;
            }
        }
        out->counts[r] = count;
    }
    {
        const double *spots = evaluation + (size_t)rows * n;
        double discount = exp(log_discounts[p->obs_nodes[rows]]);
        uint32_t count = 0;
        for (uint32_t i = 0; i < n; ++i) {
            if (!active[i]) continue;
            double payoff = call_intrinsic(spots[i], est->strike) * discount;
            if (!isfinite(payoff))
                return set_err(err, E_NON_FINITE_DISCOUNTED_PAYOFF, PH_EVALUATE, -1, i, 0, 0, 0);
            payoffs[i] = payoff;
            active[i] = 0;
            count += 1;
        }
        out->counts[rows] = count;
    }
    double mean = 0.0, m2 = 0.0;
    uint32_t count = 0;
    for (uint32_t i = 0; i < n; ++i) {
        double v = payoffs[i];
        count += 1;
        double delta = v - mean;
        mean += delta / (double)count;
        m2 = fma(delta, v - mean, m2);
    }
    out->price = mean;
    if (n > 1) {
        out->sd = sqrt(m2 / (double)(n - 1));
        out->se = out->sd / sqrt((double)n);
        out->has_sd = 1;
    } else {
        out->has_sd = 0;
        out->sd = out->se = 0.0;
    }
    return 0;
}

static void fill_result(const plan_t *p, const or_node_fit *fits, const eval_t *ev, uint32_t n,
                        or_call_result *res) {
    res->price = ev->price;
    res->paths = n;
    res->has_sd = ev->has_sd;
    res->sd = ev->sd;
    res->se = ev->se;
    res->obs = p->obs;
    for (uint32_t r = 0; r < p->obs; ++r) {
        if (res->observation_nodes) res->observation_nodes[r] = p->obs_nodes[r];
        if (res->exercise_counts) res->exercise_counts[r] = ev->counts[r];
    }
    res->fit_count = p->obs - 1;
    if (res->fits) memcpy(res->fits, fits, sizeof(or_node_fit) * (p->obs - 1));
    res->near_tie_training = 0;
    for (uint32_t r = 0; r + 1 < p->obs; ++r) res->near_tie_training += fits[r].near_ties;
    res->near_tie_evaluation = ev->near_ties;
}

int or_fit_evaluate_paths(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                          const double *training, uint32_t training_paths,
                          const double *evaluation, uint32_t evaluation_paths,
                          or_call_result *result, or_error *err) {
    plan_t p;
    plan_build(plan, &p);
    uint32_t nmax = training_paths > evaluation_paths ? training_paths : evaluation_paths;
    double *values = malloc(sizeof(double) * nmax);
    double *xb = malloc(sizeof(double) * nmax), *yb = malloc(sizeof(double) * nmax);
    uint32_t *ib = malloc(sizeof(uint32_t) * nmax);
    uint8_t *active = malloc(nmax);
    or_node_fit *fits = calloc(p.obs, sizeof(or_node_fit));
    uint32_t *counts = calloc(p.obs, sizeof(uint32_t));
    eval_t ev = {.counts = counts};
    int rc = fit_policy(&p, plan->log_discounts, initial_spot, est, training, training_paths, fits,
                        values, xb, yb, ib, err);
    if (!rc)
        rc = evaluate_policy(&p, plan->log_discounts, initial_spot, est, fits, evaluation,
                             evaluation_paths, values, active, &ev, err);
    if (!rc) fill_result(&p, fits, &ev, evaluation_paths, result);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    free(active);
    free(fits);
    free(counts);
    plan_free(&p);
    return rc;
}

/* LsmCallEstimator::price_seeded (policy.rs:306-336). */
int or_price_lsm_call(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                      uint32_t training_paths, uint32_t evaluation_paths, uint64_t training_seed,
                      uint32_t training_domain, uint64_t evaluation_seed,
                      uint32_t evaluation_domain, or_call_result *result, or_error *err) {
    if (training_domain == evaluation_domain)
        return set_err(err, E_RANDOM_DOMAIN_COLLISION, PH_NONE, -1, -1, training_domain, 0, 0);
    plan_t p;
    plan_build(plan, &p);
    int rc = 0;
    size_t obs = p.obs;
    double *train = malloc(sizeof(double) * obs * (size_t)training_paths);
    rc = simulate_population(&p, initial_spot, training_paths, training_seed, training_domain, NULL,
                             train, err, PH_TRAIN_SIM);
    uint32_t nmax = training_paths > evaluation_paths ? training_paths : evaluation_paths;
    double *values = malloc(sizeof(double) * nmax);
    double *xb = malloc(sizeof(double) * nmax), *yb = malloc(sizeof(double) * nmax);
    uint32_t *ib = malloc(sizeof(uint32_t) * nmax);
    uint8_t *active = malloc(nmax);
    or_node_fit *fits = calloc(obs, sizeof(or_node_fit));
    uint32_t *counts = calloc(obs, sizeof(uint32_t));
    eval_t ev = {.counts = counts};
    double *evalm = NULL;
    if (!rc)
        rc = fit_policy(&p, plan->log_discounts, initial_spot, est, train, training_paths, fits,
                        values, xb, yb, ib, err);
    free(train);
    if (!rc) {
        evalm = malloc(sizeof(double) * obs * (size_t)evaluation_paths);
        rc = simulate_population(&p, initial_spot, evaluation_paths, evaluation_seed,
                                 evaluation_domain, NULL, evalm, err, PH_EVAL_SIM);
    }
    if (!rc)
        rc = evaluate_policy(&p, plan->log_discounts, initial_spot, est, fits, evalm,
                             evaluation_paths, values, active, &ev, err);
    if (!rc) fill_result(&p, fits, &ev, evaluation_paths, result);
    free(evalm);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    free(active);
    free(fits);
    free(counts);
    plan_free(&p);
    return rc;
}

/* ================================ batch/pricing.rs (grouped) ================================ */

static int cmp_total(double a, double b) {
    /* f64::total_cmp for the finite values used here */
    return a < b ? -1 : (a > b ? 1 : 0);
}

/* exact_union_times (crates/finoptions-lsm/src/batch/pricing.rs:967-1047) with next_base_node
 * (:1116-1122) */
static uint32_t union_times(double maturity, uint32_t base, const double *cash, uint32_t ncash,
                            double *times) {
    uint32_t count = 0, event = 0;
    int64_t node = 0;   /* -1 once exhausted */
    while (node >= 0 || event < ncash) {
        int have_base = node >= 0;
        double bt = 0.0;
        if (have_base)
            bt = (uint32_t)node == base ? maturity
                                        : maturity * ((double)(uint32_t)node / (double)base);
        if (have_base && event < ncash) {
            int c = cmp_total(bt, cash[event]);
            if (c < 0) {
                times[count++] = bt;
                node = (uint32_t)node == base ? -1 : node + 1;
            } else if (c == 0) {
                times[count++] = bt;
                node = (uint32_t)node == base ? -1 : node + 1;
                event += 1;
            } else {
                times[count++] = cash[event];
                event += 1;
            }
        } else if (have_base) {
            times[count++] = bt;
            node = (uint32_t)node == base ? -1 : node + 1;
        } else {
            times[count++] = cash[event];
            event += 1;
        }
    }
    return count;
}

/* exact_explicit_union_times (crates/finoptions-lsm/src/batch/pricing.rs:1049-1114) */
static uint32_t explicit_union_times(double maturity, const double *ex, uint32_t nex,
                                     const double *cash, uint32_t ncash, double *times) {
    uint32_t count = 0, e = 0, c = 0;
    times[count++] = 0.0;
    while (e < nex || c < ncash) {
        double next;
        if (e < nex && c < ncash) {
            int k = cmp_total(ex[e], cash[c]);
            if (k < 0) {
                next = ex[e++];
            } else if (k == 0) {
                next = ex[e++];
                c += 1;
            } else {
                next = cash[c++];
            }
        } else if (e < nex) {
            next = ex[e++];
        } else {
            next = cash[c++];
        }
        if (next < maturity) times[count++] = next;
    }
    times[count++] = maturity;
    return count;
}

static int64_t search_time(const double *times, uint32_t n, double t) {
    uint32_t lo = 0, hi = n;
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        int c = cmp_total(times[mid], t);
        if (c == 0) return mid;
        if (c < 0) lo = mid + 1;
        else hi = mid;
    }
    return -1;
}

int or_group_grid(const or_grouped_input *in, uint32_t g, double *times, uint32_t capacity,
                  uint32_t *exercise_nodes, uint32_t *exercise_count, uint32_t *div_nodes,
                  double *div_amounts, uint32_t *div_count) {
    const double *cash = NULL, *amounts = NULL;
    uint32_t ncash = 0;
    if (in->dividend_model != 0) {   /* cash and escrowed models share the event-aligned grid */
        int64_t a = in->cash_offsets[g] - in->cash_offsets[0], b = in->cash_offsets[g + 1] - in->cash_offsets[0];
        cash = in->cash_times + a;
        amounts = in->cash_amounts + a;
        ncash = (uint32_t)(b - a);
    }
    double maturity = in->maturities[g];
    uint32_t n;
    const double *ex = NULL;
    uint32_t nex = 0;
    if (in->mesh_kind == 2) {
        int64_t a = in->exercise_offsets[g] - in->exercise_offsets[0];
        int64_t b = in->exercise_offsets[g + 1] - in->exercise_offsets[0];
        ex = in->exercise_times + a;
        nex = (uint32_t)(b - a);
        if (nex + ncash + 2 > capacity) return -1;
        n = explicit_union_times(maturity, ex, nex, cash, ncash, times);
    } else {
        uint32_t base = in->mesh_kind == 1 ? 1 : in->base_intervals;
        if (base + 1 + ncash > capacity) return -1;
        n = union_times(maturity, base, cash, ncash, times);
    }
    /* build_row_schedule (pricing.rs:893-965) */
    *div_count = ncash;
    for (uint32_t d = 0; d < ncash; ++d) {
        div_nodes[d] = (uint32_t)search_time(times, n, cash[d]);
        div_amounts[d] = amounts[d];
    }
    uint32_t ec = 0;
    if (in->mesh_kind == 2) {
        if (in->origin_exercise == 1) exercise_nodes[ec++] = 0;
        for (uint32_t e = 0; e < nex; ++e) exercise_nodes[ec++] = (uint32_t)search_time(times, n, ex[e]);
    } else {
        uint32_t start = (in->mesh_kind == 1 || in->origin_exercise == 0) ? 1 : 0;
        for (uint32_t node = start; node + 1 < n; ++node) exercise_nodes[ec++] = node;
    }
    *exercise_count = ec;
    return (int)n;
}

typedef struct {
    double *times, *logd, *q, *div_amounts;
    uint32_t *ex, *div_nodes;
    uint32_t nodes, nex, ndiv;
    or_plan_input in;
    /* escrowed model (path.rs:618-944): the plan `in` is the risky prepaid forward (no cash
     * events, zero yield); observations add prepaid[node] */
    int escrowed;
    double *prepaid;       /* [nodes] */
    double start;          /* initial simulated state: S0, or F0 = S0 - prepaid[0] */
    int prepaid_ok;        /* every prepaid value finite */
} group_plan_t;

/* path.rs:897-944 escrowed_prepaid_cash: prepaid[node] = prepaid[node + 1] exp(logD[node + 1] -
 * logD[node]) + same-node cash, expiry holding a terminal amount; prepaid[0] is then replaced by
 * the origin value sum_k D_k exp(logD[node_k]) in schedule order. Returns 0 when a value is not
 * finite. */
static int escrowed_prepaid_cash(uint32_t nodes, const double *logd, uint32_t ndiv,
                                 const uint32_t *div_nodes, const double *amounts, double *prepaid) {
    for (uint32_t i = 0; i < nodes; ++i) prepaid[i] = 0.0;
    uint32_t event = ndiv;
    const uint32_t expiry = nodes - 1;
    if (event > 0 && div_nodes[event - 1] == expiry) {
        event -= 1;
        prepaid[expiry] = amounts[event];
    }
    for (int64_t node = (int64_t)nodes - 2; node >= 0; --node) {
        double interval_log_discount = logd[node + 1] - logd[node];
        double discounted_future = prepaid[node + 1] * exp(interval_log_discount);
        double same_node_cash = 0.0;
        if (event > 0 && div_nodes[event - 1] == (uint32_t)node) {
            event -= 1;
            same_node_cash = amounts[event];
        }
        double value = discounted_future + same_node_cash;
        if (!isfinite(value)) return 0;
        prepaid[node] = value;
    }
    double origin = 0.0;
    for (uint32_t d = 0; d < ndiv; ++d) {
        origin += amounts[d] * exp(logd[div_nodes[d]]);
        if (!isfinite(origin)) return 0;
    }
    prepaid[0] = origin;
    return 1;
}

static void group_plan(const or_grouped_input *in, uint32_t g, group_plan_t *gp) {
    uint32_t ncash = 0, nex = 0;
    if (in->dividend_model != 0) ncash = (uint32_t)(in->cash_offsets[g + 1] - in->cash_offsets[g]);  /* base cancels */
    if (in->mesh_kind == 2) nex = (uint32_t)(in->exercise_offsets[g + 1] - in->exercise_offsets[g]);
    uint32_t cap = (in->mesh_kind == 0 ? in->base_intervals : 1) + nex + ncash + 3;
    gp->times = malloc(sizeof(double) * cap);
    gp->ex = malloc(sizeof(uint32_t) * cap);
    gp->div_nodes = malloc(sizeof(uint32_t) * (ncash + 1));
    gp->div_amounts = malloc(sizeof(double) * (ncash + 1));
    int n = or_group_grid(in, g, gp->times, cap, gp->ex, &gp->nex, gp->div_nodes, gp->div_amounts,
                          &gp->ndiv);
    gp->nodes = (uint32_t)n;
    gp->logd = malloc(sizeof(double) * gp->nodes);
    gp->q = malloc(sizeof(double) * gp->nodes);
    /* build_row_curves (pricing.rs:862-891) */
    double rate = in->rates[g];
    double yield = in->dividend_model == 0 ? in->dividend_yields[g] : 0.0;
    for (uint32_t i = 0; i < gp->nodes; ++i) {
        gp->logd[i] = i == 0 ? 0.0 : -rate * gp->times[i];
        gp->q[i] = i == 0 ? 0.0 : yield * gp->times[i];
    }
    gp->in.nodes = gp->nodes;
    gp->in.times = gp->times;
    gp->in.log_discounts = gp->logd;
    gp->in.cum_yields = gp->q;
    gp->in.volatility = in->volatilities[g];
    gp->in.dividend_policy = in->dividend_model == 1 ? in->cash_policy : 0;
    gp->in.dividends = gp->ndiv;
    gp->in.dividend_nodes = gp->div_nodes;
    gp->in.dividend_amounts = gp->div_amounts;
    gp->in.exercise_count = gp->nex;
    gp->in.exercise_nodes = gp->ex;
    gp->in.expiry_node = gp->nodes - 1;
    gp->escrowed = in->dividend_model == 2;
    gp->prepaid = NULL;
    gp->start = in->initial_spots[g];
    gp->prepaid_ok = 1;
    if (gp->escrowed) {
        /* EscrowedGbmPathPlan::try_new (path.rs:628-666): the risky component carries no cash
         * events and, as a prepaid forward, no yield; path_volatility (pricing.rs:840-860)
         * rescales a spot volatility by S0 / F0 with F0 from the schedule-order origin value
         * (input.rs:834-860, the same sum as prepaid[0]). */
        gp->prepaid = malloc(sizeof(double) * gp->nodes);
        gp->prepaid_ok = escrowed_prepaid_cash(gp->nodes, gp->logd, gp->ndiv, gp->div_nodes,
                                               gp->div_amounts, gp->prepaid);
        gp->in.dividends = 0;
        gp->in.dividend_policy = 0;
        double S0 = in->initial_spots[g];
        double prepaid_forward = S0 - gp->prepaid[0];
        gp->start = prepaid_forward;
        if (in->escrowed_volatility == 1)
            gp->in.volatility = in->volatilities[g] * S0 / prepaid_forward;
    }
}

static void group_plan_free(group_plan_t *gp) {
    free(gp->times);
    free(gp->ex);
    free(gp->div_nodes);
    free(gp->div_amounts);
    free(gp->logd);
    free(gp->q);
    free(gp->prepaid);
}

/* validate_escrowed_row (input.rs:798-832) and initial_prepaid_forward (path.rs:828-842). */
static int escrowed_check(const or_grouped_input *in, uint32_t g, const group_plan_t *gp,
                          or_error *err) {
    if (!gp->escrowed) return 0;
    double S0 = in->initial_spots[g];
    if (!gp->prepaid_ok)
        return set_err(err, E_ESCROWED_PREPAID_FORWARD, PH_NONE, -1, -1, g, S0, 0.0);
    double prepaid_forward = S0 - gp->prepaid[0];
    if (!isfinite(prepaid_forward) || prepaid_forward <= 0.0)
        return set_err(err, E_ESCROWED_PREPAID_FORWARD, PH_NONE, -1, -1, g, S0, gp->prepaid[0]);
    if (in->escrowed_volatility == 1) {
        double transformed = gp->in.volatility;
        if (!isfinite(transformed) || transformed <= 0.0)
            return set_err(err, E_ESCROWED_PREPAID_FORWARD, PH_NONE, -1, -1, g, transformed, 0.0);
    }
    return 0;
}

/* reconstruct_observed_stock (path.rs:844-880): origin rows observe S0; other exercise rows add
 * the prepaid cash of their node, expiry (the last grid node) adds nothing. */
static int escrowed_reconstruct(const group_plan_t *gp, const plan_t *p, double S0, double *out,
                                uint32_t paths, or_error *err, int phase) {
    for (uint32_t o = 0; o < p->obs; ++o) {
        uint32_t node = p->obs_nodes[o];
        double *row = out + (size_t)o * paths;
        if (node == 0) {
            for (uint32_t i = 0; i < paths; ++i) row[i] = S0;
            continue;
        }
        double prepaid_cash = node == gp->nodes - 1 ? 0.0 : gp->prepaid[node];
        for (uint32_t i = 0; i < paths; ++i) {
            double stock = row[i] + prepaid_cash;
            if (!isfinite(stock))
                return set_err(err, E_NON_FINITE_TRANSITION, phase, node, i, node, row[i], prepaid_cash);
            row[i] = stock;
        }
    }
    return 0;
}

/* One population of a group: the spot model directly, or the escrowed prepaid forward from F0
 * followed by the observed-stock reconstruction (path.rs:724-739). */
static int group_simulate(const group_plan_t *gp, const plan_t *p, double S0, uint32_t paths,
                          uint64_t seed, uint32_t domain, double *out, or_error *err, int phase) {
    int rc = simulate_population(p, gp->start, paths, seed, domain, NULL, out, err, phase);
    if (!rc && gp->escrowed) rc = escrowed_reconstruct(gp, p, S0, out, paths, err, phase);
    return rc;
}

/* price_group_into (pricing.rs:505-589) + try_from_parts (:1130-1230). */
static int price_group(const or_grouped_input *in, uint32_t g, or_batch_output *out,
                       or_error *err);

/* Threads of the path-parallel simulation of a single call (the reference's CpuExecutor
 * threads, runtime/src/lib.rs:364-424). */
void or_set_threads(int threads) {
#ifdef _OPENMP
    omp_set_num_threads(threads > 0 ? threads : 1);
#else
    (void)threads;
#endif
}

/* One group on the calling thread alone (simulate_group_population runs the serial simulator,
 * pricing.rs:681-701). */
static int price_group_serial(const or_grouped_input *in, uint32_t g, or_batch_output *out,
                              or_error *err) {
#ifdef _OPENMP
    const int saved = omp_get_max_threads();
    omp_set_num_threads(1);
#endif
    int rc = price_group(in, g, out, err);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return rc;
}

static int price_group(const or_grouped_input *in, uint32_t g, or_batch_output *out,
                       or_error *err) {
    group_plan_t gp;
    group_plan(in, g, &gp);
    plan_t p;
    plan_build(&gp.in, &p);
    uint32_t nt = in->training_path_counts ? in->training_path_counts[g] : in->training_paths;
    uint32_t ne = in->evaluation_path_counts ? in->evaluation_path_counts[g] : in->evaluation_paths;
    uint32_t first = in->group_offsets[g], last = in->group_offsets[g + 1];
    uint32_t strikes = last - first;
    int rc = 0;
    double *train = malloc(sizeof(double) * p.obs * (size_t)nt);
    double *evalm = NULL;
    uint32_t nmax = nt > ne ? nt : ne;
    double *values = malloc(sizeof(double) * nmax);
    double *xb = malloc(sizeof(double) * nmax), *yb = malloc(sizeof(double) * nmax);
    uint32_t *ib = malloc(sizeof(uint32_t) * nmax);
    uint8_t *active = malloc(nmax);
    or_node_fit *fits = calloc((size_t)strikes * p.obs, sizeof(or_node_fit));
    uint32_t *counts = calloc(p.obs, sizeof(uint32_t));
    uint32_t td = in->training_domain_start + g, ed = in->evaluation_domain_start + g;
    rc = escrowed_check(in, g, &gp, err);
    if (!rc)
        rc = group_simulate(&gp, &p, in->initial_spots[g], nt, in->training_seed, td, train, err,
                            PH_TRAIN_SIM);
    if (rc) {
        err->row = first;
        goto done;
    }
    for (uint32_t s = 0; s < strikes && !rc; ++s) {
        or_estimator est = {in->strikes[first + s], in->degree, in->rank_tolerance,
                            in->condition_limit, in->tie_rule, in->insufficient_support,
                            in->tie_window};
        rc = fit_policy(&p, gp.logd, in->initial_spots[g], &est, train, nt,
                        fits + (size_t)s * p.obs, values, xb, yb, ib, err);
        if (rc) err->row = first + s;
    }
    free(train);
    train = NULL;
    if (rc) goto done;
    evalm = malloc(sizeof(double) * p.obs * (size_t)ne);
    rc = group_simulate(&gp, &p, in->initial_spots[g], ne, in->evaluation_seed, ed, evalm, err,
                        PH_EVAL_SIM);
    if (rc) {
        err->row = first;
        goto done;
    }
    for (uint32_t s = 0; s < strikes; ++s) {
        uint32_t row = first + s;
        or_estimator est = {in->strikes[row], in->degree, in->rank_tolerance, in->condition_limit,
                            in->tie_rule, in->insufficient_support,
                            in->tie_window};
        eval_t ev = {.counts = counts};
        const or_node_fit *f = fits + (size_t)s * p.obs;
        rc = evaluate_policy(&p, gp.logd, in->initial_spots[g], &est, f, evalm, ne, values, active,
                             &ev, err);
        if (rc) {
            err->row = row;
            goto done;
        }
        uint8_t origin_fitted = 0, origin_exercised = 0;
        uint32_t reg = 0, nosup = 0, insuf = 0, min_itm = UINT32_MAX;
        double max_cond = 0.0;
        for (uint32_t r = 0; r + 1 < p.obs; ++r) {
            switch (f[r].kind) {
            case 0:
                origin_fitted = 1;
                origin_exercised = (uint8_t)f[r].exercise;
                break;
            case 1: nosup += 1; break;
            case 2: insuf += 1; break;
            default:
                reg += 1;
                if (f[r].itm_paths < min_itm) min_itm = f[r].itm_paths;
                if (f[r].model.condition > max_cond) max_cond = f[r].model.condition;
            }
        }
        if (reg == 0) min_itm = 0;
        uint64_t early = 0;
        for (uint32_t r = 0; r + 1 < p.obs; ++r) early += counts[r];
        out->prices[row] = ev.price;
        out->sds[row] = ev.sd;
        out->ses[row] = ev.se;
        out->origin_fitted[row] = origin_fitted;
        out->origin_exercised[row] = origin_exercised;
        out->regression_nodes[row] = reg;
        out->no_support_nodes[row] = nosup;
        out->insufficient_nodes[row] = insuf;
        out->min_itm_paths[row] = min_itm;
        out->max_condition[row] = max_cond;
        out->early_paths[row] = (uint32_t)early;
        out->expiry_paths[row] = counts[p.obs - 1];
        uint64_t train_ties = 0;
        for (uint32_t r = 0; r + 1 < p.obs; ++r) train_ties += f[r].near_ties;
        if (out->near_tie_training) out->near_tie_training[row] = train_ties;
        if (out->near_tie_evaluation) out->near_tie_evaluation[row] = ev.near_ties;
    }
done:
    free(train);
    free(evalm);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    free(active);
    free(fits);
    free(counts);
    plan_free(&p);
    group_plan_free(&gp);
    return rc;
}

/* GroupedBatchLsmCallInputs::price (pricing.rs:422-503) and BatchLsmCallInputs::price
 * (pricing.rs:191-254): serial over groups for threads <= 1, else group-parallel (a row batch is
 * a batch of one-strike groups) with every group priced serially, as the reference's private Rayon
 * pool does. The error reported is that of the lowest failing group, which is what the reference's
 * serial replay of the first failure returns. */
int or_price_grouped(const or_grouped_input *in, or_batch_output *out, int threads,
                     or_error *err) {
    zig_build();
    if (threads <= 1 || in->groups < 2) {
        for (uint32_t g = 0; g < in->groups; ++g) {
            int rc = price_group_serial(in, g, out, err);
            if (rc) return rc;
        }
        return 0;
    }
    or_error *errs = calloc(in->groups, sizeof(or_error));
    int *codes = calloc(in->groups, sizeof(int));
    if (!errs || !codes) {
        free(errs);
        free(codes);
        return set_err(err, E_INTERNAL, PH_NONE, -1, -1, 0, 0, 0);
    }
#ifdef _OPENMP
    omp_set_max_active_levels(1);
#endif
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int64_t g = 0; g < (int64_t)in->groups; ++g)
        codes[g] = price_group(in, (uint32_t)g, out, &errs[g]);
    int rc = 0;
    for (uint32_t g = 0; g < in->groups; ++g) {
        if (codes[g]) {
            rc = codes[g];
            if (err) *err = errs[g];
            break;
        }
    }
    free(errs);
    free(codes);
    return rc;
}

int or_grouped_models(const or_grouped_input *in, uint32_t g, or_node_fit *fits_out,
                      uint32_t capacity_per_strike, or_error *err) {
    zig_build();
    group_plan_t gp;
    group_plan(in, g, &gp);
    plan_t p;
    plan_build(&gp.in, &p);
    uint32_t nt = in->training_path_counts ? in->training_path_counts[g] : in->training_paths;
    uint32_t first = in->group_offsets[g], last = in->group_offsets[g + 1];
    double *train = malloc(sizeof(double) * p.obs * (size_t)nt);
    double *values = malloc(sizeof(double) * nt);
    double *xb = malloc(sizeof(double) * nt), *yb = malloc(sizeof(double) * nt);
    uint32_t *ib = malloc(sizeof(uint32_t) * nt);
    or_node_fit *fits = calloc(p.obs, sizeof(or_node_fit));
    int rc = escrowed_check(in, g, &gp, err);
    if (!rc)
        rc = group_simulate(&gp, &p, in->initial_spots[g], nt, in->training_seed,
                            in->training_domain_start + g, train, err, PH_TRAIN_SIM);
    for (uint32_t s = 0; !rc && s < last - first; ++s) {
        or_estimator est = {in->strikes[first + s], in->degree, in->rank_tolerance,
                            in->condition_limit, in->tie_rule, in->insufficient_support,
                            in->tie_window};
        rc = fit_policy(&p, gp.logd, in->initial_spots[g], &est, train, nt, fits, values, xb, yb,
                        ib, err);
        if (rc) {
            err->row = first + s;
            break;
        }
        uint32_t rows = p.obs - 1 < capacity_per_strike ? p.obs - 1 : capacity_per_strike;
        memcpy(fits_out + (size_t)s * capacity_per_strike, fits, sizeof(or_node_fit) * rows);
    }
    free(train);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    free(fits);
    plan_free(&p);
    group_plan_free(&gp);
    return rc;
}

/* Pathwise (tangent-linear) sensitivities of the frozen-policy evaluation estimator. Each path's
 * discounted payoff (S_tau - K) exp(logD(t_tau)) is differentiated with the stopping row tau held
 * fixed. Tangents of S w.r.t. (S0, sigma, r, q) follow the exact GBM transition and the dividend
 * jump (liquidator: dS_after = dS_before when S_before > D, else 0; survivor: unchanged). The rate
 * enters through logD(t) = -r t and the drift; the yield through Q(t) = q t. */
/* Test utility, no reference counterpart: the evaluation of policy.rs:642-736 applied to group g of
 * `eval_in` under the policy fitted on group g of `fit_in` (the reference's split between
 * fit_policy and evaluate, policy.rs:306-336). stop_mode 0 decides stopping rows from the frozen
 * rules, 1 does the same and records them in stops[strike * ne + path], 2 holds the recorded rows
 * fixed. prices[strike] = mean discounted payoff. Both requests must share mesh and strikes; only
 * S0, sigma, r and q may differ. */
int or_grouped_frozen_prices(const or_grouped_input *fit_in, const or_grouped_input *eval_in, uint32_t g,
                             int32_t *stops, int stop_mode, double *prices, or_error *err) {
    zig_build();
    group_plan_t gf, ge;
    group_plan(fit_in, g, &gf);
    group_plan(eval_in, g, &ge);
    plan_t pf, pe;
    plan_build(&gf.in, &pf);
    plan_build(&ge.in, &pe);
    uint32_t nt = fit_in->training_path_counts ? fit_in->training_path_counts[g] : fit_in->training_paths;
    uint32_t ne = eval_in->evaluation_path_counts ? eval_in->evaluation_path_counts[g] : eval_in->evaluation_paths;
    uint32_t first = fit_in->group_offsets[g], last = fit_in->group_offsets[g + 1];
    uint32_t strikes = last - first;
    double *train = malloc(sizeof(double) * pf.obs * (size_t)nt);
    double *values = malloc(sizeof(double) * nt);
    double *xb = malloc(sizeof(double) * nt), *yb = malloc(sizeof(double) * nt);
    uint32_t *ib = malloc(sizeof(uint32_t) * nt);
    or_node_fit *fits = calloc((size_t)strikes * pf.obs, sizeof(or_node_fit));
    double *eval = NULL;
    int rc = escrowed_check(fit_in, g, &gf, err);
    if (!rc) rc = escrowed_check(eval_in, g, &ge, err);
    if (!rc)
        rc = group_simulate(&gf, &pf, fit_in->initial_spots[g], nt, fit_in->training_seed,
                            fit_in->training_domain_start + g, train, err, PH_TRAIN_SIM);
    for (uint32_t s = 0; !rc && s < strikes; ++s) {
        or_estimator est = {fit_in->strikes[first + s], fit_in->degree, fit_in->rank_tolerance,
                            fit_in->condition_limit, fit_in->tie_rule, fit_in->insufficient_support,
                            fit_in->tie_window};
        rc = fit_policy(&pf, gf.logd, fit_in->initial_spots[g], &est, train, nt,
                        fits + (size_t)s * pf.obs, values, xb, yb, ib, err);
    }
    if (rc) goto out;
    eval = malloc(sizeof(double) * pe.obs * (size_t)ne);
    rc = group_simulate(&ge, &pe, eval_in->initial_spots[g], ne, eval_in->evaluation_seed,
                        eval_in->evaluation_domain_start + g, eval, err, PH_EVAL_SIM);
    if (rc) goto out;
    {
        const uint32_t rows = pe.obs - 1;
        const double S0 = eval_in->initial_spots[g];
        for (uint32_t sidx = 0; sidx < strikes; ++sidx) {
            const or_node_fit *f = fits + (size_t)sidx * pf.obs;
            const double K = fit_in->strikes[first + sidx];
            if (rows > 0 && f[0].kind == 0 && f[0].exercise) { /* origin exercise: deterministic */
                prices[sidx] = S0 > K ? S0 - K : 0.0;
                continue;
            }
            double sum = 0.0;
            for (uint32_t path = 0; path < ne; ++path) {
                int64_t stop = -1;
                if (stop_mode == 2) {
                    stop = stops[(size_t)sidx * ne + path];
                } else {
                    for (uint32_t r = 0; r < rows && stop < 0; ++r) {
                        if (f[r].kind != 3) continue;
                        double spot = eval[(size_t)r * ne + path];
                        if (spot <= K) continue;
                        double cont = or_predict(&f[r].model, spot / K);
                        if (should_exercise(fit_in->tie_rule, spot - K, cont)) stop = r;
                    }
                    if (stop < 0) stop = rows;
                    if (stop_mode == 1) stops[(size_t)sidx * ne + path] = (int32_t)stop;
                }
                double spot = eval[(size_t)stop * ne + path];
                double pay = spot > K ? spot - K : 0.0;
                sum += pay * exp(ge.logd[pe.obs_nodes[stop]]);
            }
            prices[sidx] = sum / ne;
        }
    }
out:
    free(train);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    free(fits);
    free(eval);
    plan_free(&pf);
    plan_free(&pe);
    group_plan_free(&gf);
    group_plan_free(&ge);
    return rc;
}

/* The tangent-linear evaluation pass shared by the grouped and the single-call Greeks: every
 * evaluation path of (seed, domain) is regenerated with its tangents, each strike's stopping row
 * is decided by its frozen fitted rules, and the discounted payoff tangents are averaged.
 * `spot_vol` is the supplied volatility (the escrowed spot-matched chain needs it),
 * `escrowed_volatility` the escrowed convention. */
static int pathwise_sums(const group_plan_t *gp, const plan_t *p, double S0, double spot_vol,
                         int escrowed_volatility, const double *strikes, uint32_t nstrikes,
                         const or_node_fit *fits, int tie_rule, uint64_t seed, uint32_t domain,
                         uint32_t ne, double *delta, double *vega, double *rho, double *dq,
                         or_error *err) {
    int rc = 0;
    const double sigma = gp->in.volatility;
    const uint32_t rows = p->obs - 1;
    double *sums = calloc((size_t)nstrikes * 4, sizeof(double));
    double *spot_obs = malloc(sizeof(double) * p->obs * 5);
    /* Escrowed chain factors. The tangents t follow the simulated state F (d/dF0, d/dsigma_F,
     * drift d/dr, d/dq). The observed stock is F + A_node(r) with F0 = S0 - A0(r) and, spot
     * matched, sigma_F = sigma S0 / F0. So dS/dS0 = t0 + t1 dsigma_F/dS0,
     * dS/dsigma = t1 dsigma_F/dsigma, dS/dr = t2 + t0 dF0/dr + t1 dsigma_F/dr + dA_node/dr.
     * The spot models use the identity factors. */
    double c_s = 0.0, c_v = 1.0, c_r0 = 0.0, c_rs = 0.0;
    double *dprepaid = calloc(gp->nodes, sizeof(double));
    if (gp->escrowed) {
        /* d/dr of the prepaid recursion: d exp(logD[n+1] - logD[n]) / dr = -(t[n+1] - t[n]) x
         * the factor; the origin value differentiates sum_k D_k exp(-r t_k). */
        for (int64_t n = (int64_t)gp->nodes - 2; n >= 0; --n) {
            double e = exp(gp->logd[n + 1] - gp->logd[n]);
            double h = gp->times[n + 1] - gp->times[n];
            dprepaid[n] = dprepaid[n + 1] * e - gp->prepaid[n + 1] * e * h;
        }
        double d0 = 0.0;
        for (uint32_t d = 0; d < gp->ndiv; ++d) {
            uint32_t node = gp->div_nodes[d];
            d0 += -gp->times[node] * gp->div_amounts[d] * exp(gp->logd[node]);
        }
        dprepaid[0] = d0;
        dprepaid[gp->nodes - 1] = 0.0;   /* expiry observes no prepaid cash */
        const double A0 = gp->prepaid[0], F0 = gp->start;
        c_r0 = -d0;
        if (escrowed_volatility == 1) {
            c_s = -spot_vol * A0 / (F0 * F0);
            c_v = S0 / F0;
            c_rs = spot_vol * S0 * d0 / (F0 * F0);
        }
    }
    for (uint32_t path = 0; path < ne && !rc; ++path) {
        stream_t s;
        stream_init(&s, seed, domain, path, p->intervals);
        double S = gp->start, t[4] = {1.0, 0.0, 0.0, 0.0};
        uint32_t observation = 0;
        if (p->obs_nodes[0] == 0) {
            spot_obs[0] = S0;
            spot_obs[1] = 1.0;
            for (int k = 1; k < 4; ++k) spot_obs[1 + k] = 0.0;
            observation = 1;
        }
        for (uint32_t iv = 0; iv < p->intervals; ++iv) {
            double z;
            if (sample_normal(&s, &z)) {
                rc = set_err(err, E_NORMAL_REJECTION_LIMIT, PH_EVAL_SIM, -1, path, iv, 0, 0);
                break;
            }
            double dt = gp->times[iv + 1] - gp->times[iv];
            double next = advance(p, iv, S, z);
            if (S == 0.0) {
                for (int k = 0; k < 4; ++k) t[k] = 0.0;
            } else {
                double growth = next / S;
                double sq = sqrt(dt);
                t[0] = t[0] * growth;
                t[1] = t[1] * growth + next * (-sigma * dt + sq * z);
                t[2] = t[2] * growth + next * dt;
                t[3] = t[3] * growth - next * dt;
            }
            S = next;
            uint32_t node = iv + 1;
            double cash = p->dividend[node];
            int has_cash = !isnan(cash);
            int terminal_cash = has_cash && node == p->obs_nodes[p->obs - 1];
            if (terminal_cash) {
                double before = S;
                S = apply_dividend(p->policy, cash, S);
                if (p->policy == 0 && !(before > cash))
                    for (int k = 0; k < 4; ++k) t[k] = 0.0;
            }
            if (observation < p->obs && p->obs_nodes[observation] == node) {
                const int expiry = node == gp->nodes - 1;
                const double add = gp->escrowed && !expiry ? gp->prepaid[node] : 0.0;
                double *o = spot_obs + observation * 5;
                o[0] = S + add;
                o[1] = t[0] + c_s * t[1];
                o[2] = c_v * t[1];
                o[3] = t[2] + c_r0 * t[0] + c_rs * t[1] + dprepaid[node];
                o[4] = t[3];
                observation += 1;
            }
            if (has_cash && !terminal_cash) {
                double before = S;
                S = apply_dividend(p->policy, cash, S);
                if (p->policy == 0 && !(before > cash))
                    for (int k = 0; k < 4; ++k) t[k] = 0.0;
            }
        }
        if (rc) break;
        for (uint32_t sidx = 0; sidx < nstrikes; ++sidx) {
            const or_node_fit *f = fits + (size_t)sidx * p->obs;
            double K = strikes[sidx];
            int64_t stop = -1;
            if (rows > 0 && f[0].kind == 0 && f[0].exercise) continue; /* deterministic */
            for (uint32_t r = 0; r < rows && stop < 0; ++r) {
                if (f[r].kind != 3) continue;
                double spot = spot_obs[r * 5];
                if (spot <= K) continue;
                double cont = or_predict(&f[r].model, spot / K);
                if (should_exercise(tie_rule, spot - K, cont)) stop = r;
            }
            uint32_t r = stop < 0 ? rows : (uint32_t)stop;
            double spot = spot_obs[r * 5];
            if (!(spot > K)) continue;
            double tn = gp->times[p->obs_nodes[r]];
            double disc = exp(gp->logd[p->obs_nodes[r]]);
            double pay = (spot - K) * disc;
            sums[sidx * 4 + 0] += spot_obs[r * 5 + 1] * disc;
            sums[sidx * 4 + 1] += spot_obs[r * 5 + 2] * disc;
            sums[sidx * 4 + 2] += spot_obs[r * 5 + 3] * disc - tn * pay;
            sums[sidx * 4 + 3] += spot_obs[r * 5 + 4] * disc;
        }
    }
    for (uint32_t sidx = 0; sidx < nstrikes && !rc; ++sidx) {
        const or_node_fit *f = fits + (size_t)sidx * p->obs;
        if (rows > 0 && f[0].kind == 0 && f[0].exercise) {
            double K = strikes[sidx];
            delta[sidx] = S0 > K ? 1.0 : 0.0;
            vega[sidx] = rho[sidx] = dq[sidx] = 0.0;
            continue;
        }
        delta[sidx] = sums[sidx * 4 + 0] / ne;
        vega[sidx] = sums[sidx * 4 + 1] / ne;
        rho[sidx] = sums[sidx * 4 + 2] / ne;
        dq[sidx] = sums[sidx * 4 + 3] / ne;
    }
    free(sums);
    free(spot_obs);
    free(dprepaid);
    return rc;
}

int or_grouped_pathwise_greeks(const or_grouped_input *in, uint32_t g, double *delta, double *vega,
                               double *rho, double *dq, or_error *err) {
    zig_build();
    group_plan_t gp;
    group_plan(in, g, &gp);
    plan_t p;
    plan_build(&gp.in, &p);
    uint32_t nt = in->training_path_counts ? in->training_path_counts[g] : in->training_paths;
    uint32_t ne = in->evaluation_path_counts ? in->evaluation_path_counts[g] : in->evaluation_paths;
    uint32_t first = in->group_offsets[g], last = in->group_offsets[g + 1];
    uint32_t strikes = last - first;
    double *train = malloc(sizeof(double) * p.obs * (size_t)nt);
    double *values = malloc(sizeof(double) * nt);
    double *xb = malloc(sizeof(double) * nt), *yb = malloc(sizeof(double) * nt);
    uint32_t *ib = malloc(sizeof(uint32_t) * nt);
    or_node_fit *fits = calloc((size_t)strikes * p.obs, sizeof(or_node_fit));
    int rc = escrowed_check(in, g, &gp, err);
    if (!rc)
        rc = group_simulate(&gp, &p, in->initial_spots[g], nt, in->training_seed,
                            in->training_domain_start + g, train, err, PH_TRAIN_SIM);
    for (uint32_t s = 0; !rc && s < strikes; ++s) {
        or_estimator est = {in->strikes[first + s], in->degree, in->rank_tolerance,
                            in->condition_limit, in->tie_rule, in->insufficient_support,
                            in->tie_window};
        rc = fit_policy(&p, gp.logd, in->initial_spots[g], &est, train, nt,
                        fits + (size_t)s * p.obs, values, xb, yb, ib, err);
    }
    free(train);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    if (!rc)
        rc = pathwise_sums(&gp, &p, in->initial_spots[g], in->volatilities[g],
                           in->escrowed_volatility, in->strikes + first, strikes, fits,
                           in->tie_rule, in->evaluation_seed, in->evaluation_domain_start + g, ne,
                           delta, vega, rho, dq, err);
    free(fits);
    plan_free(&p);
    group_plan_free(&gp);
    return rc;
}

/* Frozen-policy pathwise sensitivities of one scheduled call (the price_seeded request of
 * policy.rs:306-336, fitted as or_price_lsm_call fits it). The rate and yield directions are
 * parallel shifts of the supplied curves: logD(t) - h t and Q(t) + h t. out = delta, vega, rho,
 * dividend yield. */
int or_call_pathwise_greeks(const or_plan_input *plan, double initial_spot, const or_estimator *est,
                            uint32_t training_paths, uint32_t evaluation_paths,
                            uint64_t training_seed, uint32_t training_domain,
                            uint64_t evaluation_seed, uint32_t evaluation_domain, double *out,
                            or_error *err) {
    if (training_domain == evaluation_domain)
        return set_err(err, E_RANDOM_DOMAIN_COLLISION, PH_NONE, -1, -1, training_domain, 0, 0);
    zig_build();
    plan_t p;
    plan_build(plan, &p);
    double *train = malloc(sizeof(double) * p.obs * (size_t)training_paths);
    double *values = malloc(sizeof(double) * training_paths);
    double *xb = malloc(sizeof(double) * training_paths), *yb = malloc(sizeof(double) * training_paths);
    uint32_t *ib = malloc(sizeof(uint32_t) * training_paths);
    or_node_fit *fits = calloc(p.obs, sizeof(or_node_fit));
    int rc = simulate_population(&p, initial_spot, training_paths, training_seed, training_domain,
                                 NULL, train, err, PH_TRAIN_SIM);
    if (!rc)
        rc = fit_policy(&p, plan->log_discounts, initial_spot, est, train, training_paths, fits,
                        values, xb, yb, ib, err);
    free(train);
    free(values);
    free(xb);
    free(yb);
    free(ib);
    if (!rc) {
        group_plan_t gp;
        memset(&gp, 0, sizeof gp);
        gp.times = (double *)plan->times;
        gp.logd = (double *)plan->log_discounts;
        gp.nodes = plan->nodes;
        gp.in = *plan;
        gp.start = initial_spot;
        rc = pathwise_sums(&gp, &p, initial_spot, plan->volatility, 0, &est->strike, 1, fits,
                           est->tie_rule, evaluation_seed, evaluation_domain, evaluation_paths,
                           &out[0], &out[1], &out[2], &out[3], err);
    }
    free(fits);
    plan_free(&p);
    return rc;
}

int or_grouped_paths(const or_grouped_input *in, uint32_t g, int evaluation, double *out,
                     or_error *err) {
    zig_build();
    group_plan_t gp;
    group_plan(in, g, &gp);
    plan_t p;
    plan_build(&gp.in, &p);
    uint32_t n = evaluation
                     ? (in->evaluation_path_counts ? in->evaluation_path_counts[g] : in->evaluation_paths)
                     : (in->training_path_counts ? in->training_path_counts[g] : in->training_paths);
    int rc = escrowed_check(in, g, &gp, err);
    if (!rc)
        rc = group_simulate(&gp, &p, in->initial_spots[g], n,
                            evaluation ? in->evaluation_seed : in->training_seed,
                            (evaluation ? in->evaluation_domain_start : in->training_domain_start) + g,
                            out, err, evaluation ? PH_EVAL_SIM : PH_TRAIN_SIM);
    plan_free(&p);
    group_plan_free(&gp);
    return rc;
}

int or_escrowed_prepaid_cash(const or_grouped_input *in, uint32_t g, double *out,
                             uint32_t capacity, or_error *err) {
    group_plan_t gp;
    group_plan(in, g, &gp);
    int rc = (int)gp.nodes;
    if (!gp.escrowed) {
        rc = -set_err(err, E_INTERNAL, PH_NONE, -1, -1, g, 0.0, 0.0);
    } else if (!gp.prepaid_ok) {
        rc = -set_err(err, E_ESCROWED_PREPAID_FORWARD, PH_NONE, -1, -1, g, 0.0, 0.0);
    } else if (gp.nodes > capacity) {
        rc = -1;
    } else {
        memcpy(out, gp.prepaid, sizeof(double) * gp.nodes);
    }
    group_plan_free(&gp);
    return rc;
}
