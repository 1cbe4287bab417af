// HBM data layout shared by the host engine and the kernels.
//
// One "group" is one simulated market state (a single call is a group with one strike). Per group:
//   training paths   time-major [obs][nt] Real at path_base (coalesced per observation row)
//   values           [strikes][nt] Real at value_base: the backward-induction cashflow per strike
//   models           [strikes][obs] Model at model_base: fitted rule per (strike, observation row)
//   partials         [strikes][chunks] x kPartialStride doubles: per-chunk reductions
// Step, node and row tables are concatenated over groups and indexed by the *_base fields.
#pragma once

#include <stdint.h>

namespace fohip {

constexpr int kChunk = 1024;            // paths per reduction chunk (fixes every reduction tree)
constexpr int kPathsPerThread = 4;      // kChunk / threads per block
constexpr int kEvalChunk = 256;         // evaluation paths per chunk: one path per thread
constexpr int kBulkPathsPerThread = 8;  // fused fit pass: paths per thread
constexpr int kBulkChunk = 2048;        // fused fit pass: paths per reduction chunk
constexpr int kBulkStrikes = 8;         // fused fit pass: strikes per block
constexpr int kPartialStride = 16;      // doubles per chunk partial
constexpr int kMaxDegree = 4;

// A fused-fit chunk lies wholly inside a population of n paths (its loads need no bound). The
// product is formed in 64 bits: chunk + 1 reaches 2^21 for n near 2^32, where 32 bits wrap.
#if defined(__HIPCC__)
__host__ __device__
#endif
constexpr inline bool bulk_chunk_is_full(uint32_t chunk, uint32_t n) {
    return (static_cast<uint64_t>(chunk) + 1u) * static_cast<uint64_t>(kBulkChunk) <= n;
}

// Rule kinds of one (strike, observation row).
// kRuleNone is zero so a cleared model buffer starts with no rule.
enum RuleKind : int32_t {
    kRuleNone = 0,            // no rule (terminal row, failed fit, or not yet fitted)
    kRuleOrigin = 1,          // common valuation-origin decision
    kRuleNoSupport = 2,       // no in-the-money training state (declared Fail policy)
    kRuleContinue = 3,        // declared continuation without regression
    kRuleRegression = 4,      // fitted continuation regression
    kRulePending = 5          // regression being fitted
};

struct alignas(64) Model {
    // Evaluation form of an accepted regression, in spot units (first 72 bytes, two scalar
    // loads): continuation = hz0 + z (hz1 + z (hz2 + z (hz3 + z hz4))) with z = fma(S, a, b),
    // a = 1 / (scale K), b = -mean / scale; the same polynomial as the three-term recurrence
    // below expanded into monomials of the standardized moneyness. K is set on every fitted row.
    double hz[kMaxDegree + 1];
    double a, b;
    double K;
    int32_t kind;
    int32_t degree;
    double mean, scale;
    double alpha[kMaxDegree], beta[kMaxDegree], coef[kMaxDegree + 1], norm[kMaxDegree + 1];
    double condition;
    double prev_norm;
    double estimate;          // origin decision: estimated continuation
    uint32_t itm;             // in-the-money training paths
    int32_t rank;
    int32_t exercise;         // origin decision
    // Fused method: a regression is accepted once its coefficients exist and the rank/condition
    // gate (kGatePending) runs for every rule after the backward pass; a failed gate leaves its
    // FOHIP_E_* code and detail here for the row's error.
    int32_t gate;
    int32_t gate_index;
    double gate_observed, gate_limit;
    // Fused method: tied applications of the next row's rule counted in the pass that fits this
    // row (an integer; the row summary adds them up per strike).
    double ties;
};

constexpr int32_t kGateDone = 0;
constexpr int32_t kGatePending = -1;

struct DevGroup {
    double S0;
    double sigma;             // constant volatility (pathwise vega)
    double disc_final;        // exp(logD[first positive exercise node or expiry])
    uint64_t train_seed, eval_seed;
    uint64_t path_base;       // element offset of the training matrix
    uint64_t value_base;      // element offset of the first strike's value column
    uint32_t train_domain, eval_domain;
    uint32_t nt, ne;          // training / evaluation paths
    uint32_t intervals;       // transitions (normals per path)
    uint32_t obs;             // observation rows: exercise rows + expiry
    uint32_t origin_row;      // 1 when observation row 0 is the valuation origin
    uint32_t fit_rows;        // positive-time exercise rows = obs - 1 - origin_row
    uint32_t step_base;       // interval tables
    uint32_t node_base;       // node tables (intervals + 1 entries)
    uint32_t row_base;        // observation-row tables
    uint32_t strike_begin, strike_end;
    uint32_t model_base;      // Model index of (strike_begin, row 0)
    uint32_t chunks_t, chunks_e;
    uint32_t part_base_t;     // partial slot of (strike_begin, chunk 0), training
    uint32_t part_base_e;     // evaluation partial slot base
    uint32_t count_base;      // exercise-count slot of (strike_begin, row 0)
    int32_t policy;           // cash-dividend policy
    uint32_t chunks_b;        // fused fit chunks (kBulkChunk paths each)
    // Escrowed cash dividends (path.rs:618-944): the simulated state is the risky prepaid
    // forward F, started at start = S0 - A0, and an exercise row observes F + row_escrow[row].
    // Spot models: escrowed = 0, start = S0.
    int32_t escrowed;
    uint32_t pad1;
    double start;
    // Greek chain factors of the observed stock (identity for the spot models): dsigma_F/dS0,
    // dsigma_F/dsigma, dF0/dr, dsigma_F/dr.
    double chain[4];
    uint64_t znorm_base;      // element offset of the stored evaluation normals [intervals][ne]
};

// Node flags
constexpr uint32_t kNodeObserved = 1u;
constexpr uint32_t kNodeTerminal = 2u;
constexpr uint32_t kNodeCash = 4u;

// One interval of a group as the generation kernels read it, in one 64-byte scalar load: the
// transition coefficients of interval i (node i -> i + 1), node i + 1's flags and cash amount and,
// when node i + 1 is an observation row, that row's discount factor and time. The same values as
// the separate tables below; the kernels that step paths read this form so that their loop holds
// one table pointer instead of eight.
struct alignas(64) StepRec {
    double log_drift, diffusion, dt, sqrt_dt;
    double cash;
    double disc0, row_time;
    uint32_t flags;
    uint32_t pad;
};
static_assert(sizeof(StepRec) == 64, "one scalar load per interval");

struct DevTables {
    const DevGroup *groups;
    const StepRec *steps;         // [interval]
    const double *log_drift;      // [interval]
    const double *diffusion;      // [interval]
    const float *log_drift32;
    const float *diffusion32;
    const double *dt;             // [interval]
    const double *sqrt_dt;        // [interval]
    const double *cash;           // [node] amount (0 when absent)
    const uint32_t *node_flags;   // [node]
    const double *disc0;          // [row] exp(logD[node_row])
    const double *disc_step;      // [row] exp(logD[node_{row+1}] - logD[node_row])
    const double *row_time;       // [row] t(node_row)
    const uint32_t *row_node;     // [row]
    const double *strikes;        // [strike row]
    const uint32_t *strike_group; // [strike row]
    const double *row_escrow;     // [observation row] prepaid cash observed (escrowed groups)
    const double *row_escrow_dr;  // [observation row] its derivative w.r.t. the rate
    const double *inv_strikes;    // [strike row] 1 / K
};

// One entry of the fit sweep's schedule (single-strike groups): the block that takes it joins the
// team sweeping `group` and keeps the group's fused-fit chunks [chunk_begin, chunk_end) for the
// whole backward induction. Blocks take entries by ticket, one team's entries after another.
struct SweepEntry {
    uint32_t group;
    uint32_t chunk_begin, chunk_end;
    uint32_t members;      // blocks in the team
    uint32_t team;         // the team's arrival counter
    uint32_t leader;       // 1: this member stores the team's rules and failures
    uint64_t part_base;    // first partial record of the team: 2 step parities x chunks_b records
};
static_assert(sizeof(SweepEntry) == 32, "two entries per 64-byte line");
constexpr int kSweepStride = 18;   // doubles per sweep partial record: 14 sums, pivot, ties, failure, pad
constexpr int kSweepChunks64 = 3;  // chunks a sweep block keeps in registers (binary64 storage)
constexpr int kSweepChunks32 = 5;  // the same with binary32 storage

// Per-strike-row failure record: the smallest key wins (earliest failure in processing order).
struct RowError {
    unsigned long long key;       // (step << 44) | (stage << 36) | path  (path < 2^36)
    int32_t code;
    int32_t pad;
    double observed, limit;
    int64_t index;
};

constexpr unsigned long long kNoError = ~0ull;
constexpr uint32_t kEvalStepBase = 0x80000u;   // evaluation failures: step field = base + row

struct RowSummary {
    double max_condition;
    double origin_estimate;
    double ties_training;   // sum of the rules' `ties` (fused method)
    uint32_t regression_nodes, no_support_nodes, insufficient_nodes, min_itm;
    uint32_t early, expiry;
    int32_t nosupport_row;
    uint32_t nosupport_count;
    uint8_t origin_fitted, origin_exercised;
    uint8_t pad[6];
};

__host__ __device__ inline unsigned long long error_key(uint32_t step, uint32_t stage, uint64_t path) {
    return (static_cast<unsigned long long>(step) << 44) | (static_cast<unsigned long long>(stage) << 36) |
           (path & ((1ull << 36) - 1));
}

// stages inside one backward step, in the reference's processing order
enum Stage : uint32_t {
    kStageDiscount = 1,       // discount_values: non-finite cashflow
    kStagePredictor = 2,      // non-finite moneyness
    kStageRegression = 3,     // regression / support failure (reduce kernels)
    kStageApply = 4,          // training-rule prediction failure
    kStageEvalPrediction = 5, // evaluation: non-finite continuation prediction
    kStageEvalPayoff = 6      // evaluation: non-finite discounted payoff
};

}  // namespace fohip
