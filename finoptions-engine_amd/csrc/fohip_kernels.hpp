// Host-callable launchers of the LSM kernels (implemented in fohip_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "fohip.h"
#include "fohip_layout.hpp"

namespace fohip {

struct FitParams {
    int32_t degree;
    int32_t tie;        // 0 strictly greater, 1 greater or equal
    int32_t support;    // 0 fail, 1 continue
    int32_t pad;
    double rank_tol;
    double cond_limit;
    // near-tie window of exercise decisions: |intrinsic - continuation| <= tie_window |continuation|
    double tie_window;
};

// Schedule and team state of one fit sweep (see k_fit_sweep).
struct SweepPlan {
    const SweepEntry *entries;    // in ticket order: team by team, each team's members in turn
    uint32_t *sync;               // [teams] arrival counters, the abort word, the ticket counter (all
                                  // zero before the launch)
    double *partials;             // kSweepStride doubles per record
    uint32_t entry_count;
    uint32_t teams;
    uint32_t spin_limit;          // polls of one wait before the sweep gives up and aborts
    uint32_t pad;
};

// Word fills of up to kMaxFills device regions in one launch (the per-pricing initialisation of
// rules, counters and failure keys); each region is `words` 32-bit words set to `value`.
constexpr int kMaxFills = 8;
struct FillRegion {
    uint32_t *ptr;
    uint64_t words;
    uint32_t value;
    uint32_t pad;
};
struct FillList {
    FillRegion region[kMaxFills];
    int32_t count;
};
hipError_t launch_fill(hipStream_t st, const FillList &fills);

hipError_t launch_philox_blocks(hipStream_t st, uint64_t seed, uint32_t domain, uint32_t path0,
                                uint32_t paths, uint32_t blocks, uint32_t *out);
hipError_t launch_gbm_exp(hipStream_t st, const double *x, double *out, uint32_t n);
hipError_t launch_stream_normals(hipStream_t st, uint64_t seed, uint32_t domain, uint32_t path0,
                                 uint32_t paths, uint32_t count, double *out, uint32_t *draws,
                                 int *fail);
hipError_t launch_simulate(hipStream_t st, bool f32, const DevTables &t, const uint2 *items,
                           uint32_t nitems, void *paths, int evaluation, unsigned long long *fail);
// FP64 method: prepare (apply + discount + Welford), then per degree k the moments of p_k
hipError_t launch_fit_prepare(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step,
                              const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                              void *values, const Model *models, double *partials, RowError *errors,
                              unsigned long long *ties);
hipError_t launch_fit_reduce_welford(hipStream_t st, const DevTables &t, const FitParams &fp,
                                     uint32_t step, const uint32_t *rows, uint32_t nrows,
                                     double *partials, Model *models, RowError *errors);
hipError_t launch_fit_moments(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step, int k,
                              const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                              const void *values, const Model *models, double *partials, const RowError *errors);
hipError_t launch_fit_reduce_moments(hipStream_t st, const DevTables &t, const FitParams &fp,
                                     uint32_t step, int k, const uint32_t *rows, uint32_t nrows,
                                     double *partials, Model *models, RowError *errors);
// Fused method: one HBM pass per backward step (apply, Welford + Gram via pivoted power sums) over
// (group, kBulkChunk-path chunk) items, then one wave per strike row reduces and fits.
// `stops` non-null selects the stopping-row encoding of the cashflows (one byte per strike and path,
// obs <= 255) instead of the value columns `values`; both give bitwise the same sums.
hipError_t launch_fit_bulk_step(hipStream_t st, bool f32, const DevTables &t, const FitParams &fp, uint32_t step,
                                const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                                void *values, uint8_t *stops, const Model *models, double *partials,
                                RowError *errors);
hipError_t launch_fit_bulk_reduce(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step,
                                  const uint32_t *rows, uint32_t nrows, double *partials, Model *models,
                                  RowError *errors);
// Fused method, single-strike groups: the whole backward induction of every group of the batch in
// one persistent launch of `slots` blocks that take the plan's entries by ticket (bitwise the
// per-node launches' rules, failures and final cashflows). Any grid is safe once the device can
// hold the largest team's blocks at once.
hipError_t launch_fit_sweep(hipStream_t st, bool f32, const DevTables &t, const FitParams &fp,
                            const SweepPlan &plan, uint32_t slots, const void *paths, void *values,
                            Model *models, RowError *errors);
// Blocks of the sweep kernel one CU holds at once (0 on failure).
int sweep_resident_blocks(bool f32, bool window);
hipError_t launch_fit_final(hipStream_t st, bool f32, bool fused, const DevTables &t, const FitParams &fp,
                            const uint2 *items, uint32_t nitems, uint32_t max_strikes,
                            const void *paths, const void *values, const uint8_t *stops, const Model *models,
                            double *partials, RowError *errors, unsigned long long *ties);
hipError_t launch_fit_reduce_origin(hipStream_t st, const DevTables &t, const FitParams &fp,
                                    const uint32_t *rows, uint32_t nrows, double *partials,
                                    Model *models, RowError *errors);
hipError_t launch_evaluate(hipStream_t st, bool fast, uint32_t greek_mask, bool escrow, uint32_t tile,
                           const DevTables &t, const FitParams &fp, const uint2 *items,
                           uint32_t nitems, uint32_t max_strikes, uint32_t max_obs,
                           const Model *models, double *partials, uint32_t *counts,
                           uint32_t *nosupport, RowError *errors, unsigned long long *sim_fail,
                           const double *znorm);
// Deferred rank/condition gate of every fused rule, then each row's earliest gate failure.
hipError_t launch_fit_gate(hipStream_t st, const DevTables &t, const FitParams &fp, const uint32_t *rows,
                           uint32_t nrows, Model *models, uint32_t model_count, RowError *errors);
// Evaluation normals [intervals][ne] per group at DevGroup::znorm_base (one block per kEvalChunk paths).
hipError_t launch_eval_normals(hipStream_t st, const DevTables &t, const uint2 *items, uint32_t nitems,
                               double *znorm, unsigned long long *sim_fail);
hipError_t launch_eval_reduce(hipStream_t st, const DevTables &t, const uint32_t *rows,
                              uint32_t nrows, const double *partials, int greeks, double *out);

hipError_t launch_row_summary(hipStream_t st, const DevTables &t, const Model *models,
                              const uint32_t *counts, const uint32_t *nosupport, uint32_t rows,
                              RowSummary *out);
hipError_t launch_replay_path(hipStream_t st, const DevTables &t, uint32_t group, int evaluation,
                              uint32_t path, double *out);

// Device region of the allocation-free entry point: `bytes` of a file-scope device array, its
// address on `device` (which must be current) taken once and kept per device.
constexpr size_t kNoAllocScratchBytes = FOHIP_NO_ALLOC_SCRATCH_BYTES;
constexpr int kNoAllocMaxDevices = 64;
hipError_t noalloc_scratch(int device, void **region, size_t *bytes);

// Strike tiles available to the evaluation kernel for a given Greeks setting.
uint32_t eval_tile_for(uint32_t max_strikes, bool greeks, uint32_t requested);

}  // namespace fohip
