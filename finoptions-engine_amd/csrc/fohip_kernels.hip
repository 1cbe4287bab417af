// CDNA4 kernels of the LSM hot path: seeded simulation, backward-induction fit passes,
// deterministic chunk reductions, and the fused regenerate-and-evaluate pass (with pathwise Greeks).
//
// Work decomposition. A training work item is (group, 1024-path chunk); fit passes launch one
// 256-thread block per (work item, strike) and every thread owns paths
// chunk*1024 + threadIdx.x + 256*k, k = 0..3, so HBM rows are read fully coalesced. The block
// index is mapped XCD-aware: the MI355X dispatcher deals workgroups round-robin over its 8 XCDs,
// so every strike of one chunk is given a block id with the same residue mod 8 and the chunk's
// spot rows are fetched into one XCD's L2 once and reused by all of its strikes.
// Evaluation work items are (group, 256-path chunk) with one path per thread.
// Reductions go thread -> wave (down-shuffle tree) -> block (fixed order) -> chunk list (fixed
// order in a one-block-per-strike reduce kernel): results are independent of timing and tuning.
#include <hip/hip_runtime.h>

#include <mutex>
#include <type_traits>

#include "fohip_device.hpp"
#include "fohip.h"
#include "fohip_kernels.hpp"

namespace fohip {

#include "zig_tables.inc"

__device__ const double g_zig_x[257] = FOHIP_ZIG_X_INIT;
__device__ const ExpEntry g_exp_table[32] = FOHIP_EXP_TABLE_INIT;
// F is read only by the exact wedge test inside the squeeze's band (about 1e-5 of the normals),
// straight from global memory; the layer abscissae and the squeeze live in LDS.
__device__ const double g_zig_f[257] = FOHIP_ZIG_F_INIT;
__device__ const float4 g_zig_squeeze[256] = FOHIP_ZIG_SQUEEZE_INIT;

__device__ __forceinline__ void load_zig(double *sx, float4 *ssq) {
    for (int i = threadIdx.x; i < 257; i += blockDim.x) sx[i] = g_zig_x[i];
    for (int i = threadIdx.x; i < 256; i += blockDim.x) ssq[i] = g_zig_squeeze[i];
    __syncthreads();
}

// the ZIGNOR tables and the exponential's table of a kernel that steps paths
__device__ __forceinline__ void load_zig_exp(double *sx, float4 *ssq, ExpEntry *se) {
    for (int i = threadIdx.x; i < 32; i += blockDim.x) se[i] = g_exp_table[i];
    load_zig(sx, ssq);
}

// Path-level failures only lower the row's key; the failure kind is a function of the key's stage
// and is decoded on the host, so racing writers cannot leave a key/code mismatch.
__device__ __forceinline__ void record_error(RowError *e, unsigned long long key) { atomicMin(&e->key, key); }

// XCD-aware decomposition of a 1-D grid into (work item, column): block b runs on XCD b mod 8;
// item = 8 * ((b / 8) / columns) + b mod 8 and column = (b / 8) mod columns, so all columns of an
// item share an XCD and run back to back there.
__device__ __forceinline__ bool xcd_item_column(uint32_t nitems, uint32_t columns, uint32_t &item,
                                                uint32_t &column) {
    const uint32_t b = blockIdx.x;
    const uint32_t k = b >> 3;
    column = k % columns;
    item = (k / columns) * 8u + (b & 7u);
    return item < nitems;
}

// Backward steps alternate the direction in which a launch walks its work items: odd steps take
// them last to first. Blocks are dispatched roughly in block order, so the first blocks of a step
// read the spot rows and cashflow columns that the previous step read last, while those lines are
// still held by the 256 MB memory-side Infinity Cache (about a quarter of a large step's reads
// then come from it instead of HBM). Every item keeps its own partial slots, so results do not
// depend on the direction.
__device__ __forceinline__ uint32_t serpentine_item(uint32_t step, uint32_t nitems, uint32_t wi) {
    return (step & 1u) ? nitems - 1u - wi : wi;
}

// Work item wi as wave-uniform scalars.
__device__ __forceinline__ uint2 load_item(const uint2 *items, uint32_t wi) {
    const uint32_t *w = reinterpret_cast<const uint32_t *>(items + wi);
    return make_uint2(*uniform(w), *uniform(w + 1));
}

// ------------------------------------------------------------------------ random-stream probes

__global__ void k_philox_blocks(uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths, uint32_t blocks,
                                uint32_t *out) {
    const uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= static_cast<uint64_t>(paths) * blocks) return;
    const uint32_t p = static_cast<uint32_t>(i / blocks), b = static_cast<uint32_t>(i % blocks);
    const uint4 w = philox4x32_10(b, 0u, path0 + p, domain, static_cast<uint32_t>(seed),
                                  static_cast<uint32_t>(seed >> 32));
    out[i * 4 + 0] = w.x;
    out[i * 4 + 1] = w.y;
    out[i * 4 + 2] = w.z;
    out[i * 4 + 3] = w.w;
}

__global__ void k_stream_normals(uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths, uint32_t count,
                                 double *out, uint32_t *draws, int *fail) {
    __shared__ double sx[257];
    __shared__ float4 ssq[256];
    __shared__ ExpEntry se[32];
    __shared__ uint64_t ring[kRingSlots * kBlock];
    load_zig_exp(sx, ssq, se);
    const uint32_t p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= paths) return;
    RingStream s;
    s.init(seed, domain, path0 + p, ring, threadIdx.x, kBlock);
    for (uint32_t i = 0; i < count; ++i) {
        double z = 0.0;
        uint32_t d = 0;
        s.top_up(true);
        if (!zignor_normal(s, sx, ssq, g_zig_f, FOHIP_ZIG_R, z, d, se)) atomicExch(fail, 1);
        out[static_cast<uint64_t>(p) * count + i] = z;
        draws[static_cast<uint64_t>(p) * count + i] = d;
    }
}

// ------------------------------------------------------------------------------ simulation

// seeded.rs:399-459: one thread per path; for each interval draw a normal, advance, then apply a
// terminal cash event before recording expiry, or record a pre-event exercise state before a
// same-node cash event. The state evolves in binary64 for both methods; the FP32-bulk method
// stores the observations in binary32 (a random-walk of binary32 transitions accumulates ~7e-7
// relative error over 50 steps, which moves LSM stopping decisions; storage rounding does not).
template <typename Store>
__global__ __launch_bounds__(kBlock) void k_simulate(DevTables t, const uint2 *__restrict__ items,
                                                     Store *__restrict__ paths_out, int evaluation,
                                                     unsigned long long *__restrict__ fail) {
    __shared__ double sx[257];
    __shared__ float4 ssq[256];
    __shared__ ExpEntry se[32];
    __shared__ uint64_t ring[kRingSlots * kBlock];
    load_zig_exp(sx, ssq, se);
    const uint2 item = items[blockIdx.x];
    // the group's fields as wave-uniform values read once: through a plain reference the compiler
    // reloads them after every store of the path loop
    const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + item.x);
    const uint32_t n = evaluation ? g->ne : g->nt;
    const uint32_t p = item.y * kChunk + blockIdx.y * kBlock + threadIdx.x;
    if (p >= n) return;
    RingStream s;
    s.init(evaluation ? g->eval_seed : g->train_seed, evaluation ? g->eval_domain : g->train_domain, p, ring,
           threadIdx.x, kBlock);
    const uint32_t intervals = g->intervals, row_base = g->row_base;
    const int policy = g->policy;
    const bool escrowed = g->escrowed != 0;
    const FOHIP_UNIFORM StepRec *sr = uniform(t.steps + g->step_base);
    Store *out = paths_out + g->path_base + p;
    double spot = g->start;
    uint32_t row = 0;
    if (g->origin_row) {
        __builtin_nontemporal_store(static_cast<Store>(g->S0), out);
        row = 1;
    }
    bool ok = true;
    for (uint32_t i = 0; i < intervals; ++i, ++sr) {
        // the interval's record, asked for before the normal is drawn so that the scalar loads
        // complete while it is
        const double diffusion = sr->diffusion, log_drift = sr->log_drift, amount = sr->cash;
        const uint32_t flags = sr->flags;
        double z;
        uint32_t draws = 0;
        s.top_up(true);
        if (!zignor_normal(s, sx, ssq, g_zig_f, FOHIP_ZIG_R, z, draws, se)) {
            ok = false;
            break;
        }
        double growth;
        const double next = gbm_advance(spot, diffusion, log_drift, z, growth, se);
        if (!isfinite(next)) {
            ok = false;
            break;
        }
        spot = next;
        const bool cash = flags & kNodeCash;
        const bool terminal = flags & kNodeTerminal;
        if (cash && terminal) spot = apply_cash<double>(policy, amount, spot);
        if (flags & kNodeObserved) {
            // escrowed groups observe F + A(node) (path.rs:844-880); A is 0 at expiry
            const double observed = escrowed ? spot + t.row_escrow[row_base + row] : spot;
            // a streaming store: the matrix is read back only by the backward pass, rows in reverse
            __builtin_nontemporal_store(static_cast<Store>(observed), out + static_cast<uint64_t>(row) * n);
            row += 1;
        }
        if (cash && !terminal) spot = apply_cash<double>(policy, amount, spot);
    }
    if (!ok) atomicMin(fail + item.x, static_cast<unsigned long long>(p));
}

// ------------------------------------------------------------------------------- fills

// blockIdx.y selects the region; the x blocks stride over its words
__global__ __launch_bounds__(kBlock) void k_fill(FillList fills) {
    const FillRegion f = fills.region[blockIdx.y];
    for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * kBlock + threadIdx.x; i < f.words;
         i += static_cast<uint64_t>(gridDim.x) * kBlock)
        f.ptr[i] = f.value;
}

hipError_t launch_fill(hipStream_t st, const FillList &fills) {
    if (fills.count <= 0) return hipSuccess;
    if (fills.count > kMaxFills) return hipErrorInvalidValue;
    uint64_t most = 0;
    for (int i = 0; i < fills.count; ++i) most = most > fills.region[i].words ? most : fills.region[i].words;
    if (most == 0) return hipSuccess;
    const uint64_t blocks = (most + kBlock - 1) / kBlock;
    const dim3 grid(static_cast<uint32_t>(blocks < 512 ? blocks : 512), static_cast<uint32_t>(fills.count));
    hipLaunchKernelGGL(k_fill, grid, dim3(kBlock), 0, st, fills);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------- fit passes

__device__ __forceinline__ bool row_dead(const RowError *err, uint32_t s) { return err[s].key != kNoError; }

// A rule application is a tie when intrinsic == continuation: the tie rule settles it, and the
// other rule would decide it the other way. With a window w > 0 the counts take every near tie,
// |intrinsic - continuation| <= w |continuation|, the decisions a last-bit difference in the
// continuation (another summation order, another form of the same polynomial) could flip.
// Counted per strike row, training and evaluation apart, and reported with the results. The
// exact test is the decision's own equality comparison; the windowed one is a diagnostic form
// (template argument WIN of the fused fit step and of the evaluation).
__device__ __forceinline__ bool near_tie(double intrinsic, double continuation, double window) {
    return fabs(intrinsic - continuation) <= window * fabs(continuation);
}

__device__ __forceinline__ bool is_tie(double intrinsic, double continuation, double window) {
    return window == 0.0 ? intrinsic == continuation : near_tie(intrinsic, continuation, window);
}

// Rare path: near ties are counted with one atomic per wave (or lane) that has any.
__device__ __noinline__ void add_ties(unsigned long long *counter, uint32_t c) {
    atomicAdd(counter, static_cast<unsigned long long>(c));
}

template <typename Real>
__device__ __forceinline__ Real intrinsic(Real spot, Real strike) {
    const Real v = spot - strike;
    return v > Real(0) ? v : Real(0);
}

__device__ __forceinline__ double *partial_slot(double *partials, const DevGroup &g, uint32_t local, uint32_t chunk) {
    return partials + static_cast<uint64_t>(g.part_base_t + local * g.chunks_t + chunk) * kPartialStride;
}

// FP64 step j of backward induction over exercise row r = obs - 2 - j:
//   j == 0: cashflow = terminal intrinsic; j > 0: apply the fitted rule of row r + 1
//   (policy.rs:570-601), then discount from node(r + 1) to node(r) (policy.rs:1049-1063),
//   store, and accumulate Welford moments of moneyness S/K over in-the-money paths
//   (regression.rs:279-297).
__global__ __launch_bounds__(kBlock) void k_fit_prepare(DevTables t, FitParams fp, uint32_t step,
                                                        const uint2 *__restrict__ items, uint32_t nitems,
                                                        uint32_t columns, const double *__restrict__ paths,
                                                        double *__restrict__ values, const Model *__restrict__ models,
                                                        double *__restrict__ partials, RowError *__restrict__ errors,
                                                        unsigned long long *__restrict__ ties) {
    __shared__ double scratch[3 * kWaves];
    uint32_t wi, local;
    if (!xcd_item_column(nitems, columns, wi, local)) return;
    const uint2 item = items[wi];
    const DevGroup &g = t.groups[item.x];
    if (g.strike_begin + local >= g.strike_end || step >= g.fit_rows) return;
    const uint32_t srow = g.strike_begin + local;
    if (row_dead(errors, srow)) return;
    const uint32_t r = g.obs - 2 - step;
    const double K = t.strikes[srow];
    const uint32_t n = g.nt;
    const double *S_r = paths + g.path_base + static_cast<uint64_t>(r) * n;
    const double *S_next = paths + g.path_base + static_cast<uint64_t>(r + 1) * n;
    double *V = values + g.value_base + static_cast<uint64_t>(local) * n;
    const Model &m_next = models[g.model_base + local * g.obs + r + 1];
    const bool apply = step > 0 && m_next.kind == kRuleRegression;
    const double disc = t.disc_step[g.row_base + r];
    Moments acc = {0.0, 0.0, 0.0};
    uint32_t tie_count = 0;
#pragma unroll
    for (int k = 0; k < kPathsPerThread; ++k) {
        const uint32_t p = item.y * kChunk + k * kBlock + threadIdx.x;
        if (p >= n) break;
        double v;
        if (step == 0) {
            v = intrinsic(S_next[p], K);
        } else {
            v = V[p];
            if (apply) {
                const double s1 = S_next[p];
                if (s1 > K) {
                    const double cont = predict(m_next, fp.degree, s1 / K);
                    if (!isfinite(cont)) {
                        record_error(&errors[srow], error_key(step - 1, kStageApply, p));
                    } else {
                        tie_count += is_tie(s1 - K, cont, fp.tie_window) ? 1u : 0u;
                        if (should_exercise(fp.tie, s1 - K, cont)) v = s1 - K;
                    }
                }
            }
        }
        v *= disc;
        if (!isfinite(v)) record_error(&errors[srow], error_key(step, kStageDiscount, p));
        V[p] = v;
        const double s0 = S_r[p];
        if (s0 > K) {
            const double x = s0 / K;
            if (!isfinite(x)) record_error(&errors[srow], error_key(step, kStagePredictor, p));
            welford_push(acc, x);
        }
    }
    if (tie_count != 0u) add_ties(ties + srow, tie_count);
    acc = block_chan(acc, scratch);
    if (threadIdx.x == 0) {
        double *out = partial_slot(partials, g, local, item.y);
        out[0] = acc.n;
        out[1] = acc.mean;
        out[2] = acc.m2;
    }
}

__device__ void fail_row(RowError *e, uint32_t step, int32_t code, int64_t index, double observed, double limit) {
    const unsigned long long key = error_key(step, kStageRegression, 0);
    if (key < e->key) {
        e->key = key;
        e->code = code;
        e->index = index;
        e->observed = observed;
        e->limit = limit;
    }
}

// Support decision of one (strike, row) from the in-the-money count (policy.rs:480-529). Returns
// true when a regression is to be fitted.
__device__ bool decide_support(Model &m, const FitParams &fp, RowError *e, uint32_t step, uint32_t count,
                               double strike) {
    const uint32_t required = static_cast<uint32_t>(fp.degree) + 1;
    m.K = strike;
    m.itm = count;
    m.rank = 0;
    m.condition = 0.0;
    if (count == 0) {
        m.kind = fp.support == 1 ? kRuleContinue : kRuleNoSupport;
        return false;
    }
    if (count < required) {
        if (fp.support == 1) {
            m.kind = kRuleContinue;
        } else {
            m.kind = kRuleNone;
            fail_row(e, step, FOHIP_E_INSUFFICIENT_OBSERVATIONS, count, required, fp.degree);
        }
        return false;
    }
    return true;
}

__device__ void begin_fit(Model &m, double mean, double scale) {
    m.mean = mean;
    m.scale = scale;
    m.prev_norm = 1.0;
#pragma unroll
    for (int k = 0; k < kMaxDegree; ++k) m.alpha[k] = m.beta[k] = 0.0;
#pragma unroll
    for (int k = 0; k <= kMaxDegree; ++k) m.coef[k] = m.norm[k] = 0.0;
    m.kind = kRulePending;
}

// Predictor location/scale of one (strike, row) (regression.rs:315-325).
__global__ __launch_bounds__(kBlock) void k_fit_reduce_welford(DevTables t, FitParams fp, uint32_t step,
                                                               const uint32_t *__restrict__ rows,
                                                               double *__restrict__ partials,
                                                               Model *__restrict__ models,
                                                               RowError *__restrict__ errors) {
    __shared__ double scratch[3 * kWaves];
    const uint32_t srow = rows[blockIdx.x];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    if (step >= g.fit_rows || row_dead(errors, srow)) return;
    const uint32_t local = srow - g.strike_begin;
    const uint32_t r = g.obs - 2 - step;
    Moments acc = {0.0, 0.0, 0.0};
    for (uint32_t c = threadIdx.x; c < g.chunks_t; c += kBlock) {
        const double *in = partial_slot(partials, g, local, c);
        acc = chan_merge(acc, Moments{in[0], in[1], in[2]});
    }
    acc = block_chan(acc, scratch);
    if (threadIdx.x != 0) return;
    Model &m = models[g.model_base + local * g.obs + r];
    if (!decide_support(m, fp, &errors[srow], step, static_cast<uint32_t>(acc.n), t.strikes[srow])) return;
    double mean = 0.0, scale = 1.0;
    if (fp.degree != 0) {
        scale = sqrt(acc.m2 / acc.n);
        if (!isfinite(acc.mean) || !isfinite(scale) || scale <= 0.0) {
            m.kind = kRuleNone;
            fail_row(&errors[srow], step, FOHIP_E_DEGENERATE_PREDICTOR, fp.degree, scale, 0.0);
            return;
        }
        mean = acc.mean;
    }
    begin_fit(m, mean, scale);
}

// Moments of the basis vector p_k (runtime/src/lib.rs:84-95, 141-158): p_k is rebuilt from the
// standardized predictor with the recurrence coefficients fitted so far, so no basis vector is
// stored between passes.
__global__ __launch_bounds__(kBlock) void k_fit_moments(DevTables t, FitParams fp, uint32_t step, int k,
                                                        const uint2 *__restrict__ items, uint32_t nitems,
                                                        uint32_t columns, const double *__restrict__ paths,
                                                        const double *__restrict__ values,
                                                        const Model *__restrict__ models, double *__restrict__ partials,
                                                        const RowError *__restrict__ errors) {
    __shared__ double scratch[3 * kWaves];
    uint32_t wi, local;
    if (!xcd_item_column(nitems, columns, wi, local)) return;
    const uint2 item = items[wi];
    const DevGroup &g = t.groups[item.x];
    if (g.strike_begin + local >= g.strike_end || step >= g.fit_rows) return;
    const uint32_t srow = g.strike_begin + local;
    const uint32_t r = g.obs - 2 - step;
    const Model &m = models[g.model_base + local * g.obs + r];
    if (row_dead(errors, srow) || m.kind != kRulePending) return;
    const double K = t.strikes[srow];
    const uint32_t n = g.nt;
    const double *S_r = paths + g.path_base + static_cast<uint64_t>(r) * n;
    const double *V = values + g.value_base + static_cast<uint64_t>(local) * n;
    double sums[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int q = 0; q < kPathsPerThread; ++q) {
        const uint32_t p = item.y * kChunk + q * kBlock + threadIdx.x;
        if (p >= n) break;
        const double s0 = S_r[p];
        if (!(s0 > K)) continue;
        const double x = s0 / K;
        const double z = fp.degree == 0 ? 0.0 : (x - m.mean) / m.scale;
        const double y = V[p];
        double previous = 0.0, current = 1.0;
        for (int j = 0; j < k; ++j) {
            const double next = fma(z - m.alpha[j], current, -m.beta[j] * previous);
            previous = current;
            current = next;
        }
        const double sq = current * current;
        sums[0] += sq;
        sums[1] = fma(y, current, sums[1]);
        sums[2] = fma(z, sq, sums[2]);
    }
    block_sum<3>(sums, scratch);
    if (threadIdx.x == 0) {
        double *out = partial_slot(partials, g, local, item.y);
        out[0] = sums[0];
        out[1] = sums[1];
        out[2] = sums[2];
    }
}

// ---- 5x5 conditioning diagnostic (regression.rs:540-577, 711-850), one thread per fit.
// Loops run over the fixed maximum of five columns with the degree as a guard, so every array
// stays in registers.

__device__ int orthonormal_transform(int degree, const Model &m, double (&tr)[25]) {
    double previous[5] = {0, 0, 0, 0, 0}, current[5] = {1, 0, 0, 0, 0};
#pragma unroll
    for (int i = 0; i < 25; ++i) tr[i] = 0.0;
#pragma unroll
    for (int k = 0; k <= kMaxDegree; ++k) {
        if (k > degree) break;
        const double inverse_norm = 1.0 / sqrt(m.norm[k]);
        if (!isfinite(inverse_norm)) return k + 1;
#pragma unroll
        for (int j = 0; j <= k; ++j) tr[k * 5 + j] = current[j] * inverse_norm;
        if (k == degree) break;
        double next[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j <= k; ++j) {
            next[j + 1] += current[j];
            next[j] -= m.alpha[k] * current[j];
            next[j] -= m.beta[k] * previous[j];
        }
#pragma unroll
        for (int j = 0; j < 5; ++j)
            if (!isfinite(next[j])) return k + 2;
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            previous[j] = current[j];
            current[j] = next[j];
        }
    }
    return 0;
}

template <int COLUMNS>
__device__ int jacobi_singular_values(double (&mat)[25], double (&sv)[5]) {
    const double tol = 16.0 * kEpsilon;
    bool converged = false;
    for (int sweep = 0; sweep < 32; ++sweep) {
        converged = true;
#pragma unroll
        for (int left = 0; left < COLUMNS - 1; ++left) {
#pragma unroll
            for (int right = left + 1; right < COLUMNS; ++right) {
                double ln = 0.0, rn = 0.0, cross = 0.0;
#pragma unroll
                for (int row = 0; row < COLUMNS; ++row) {
                    const double lv = mat[left * 5 + row], rv = mat[right * 5 + row];
                    ln = fma(lv, lv, ln);
                    rn = fma(rv, rv, rn);
                    cross = fma(lv, rv, cross);
                }
                if (!isfinite(ln) || !isfinite(rn) || !isfinite(cross) || ln <= 0.0 || rn <= 0.0) return 1 + right;
                if (fabs(cross) <= tol * sqrt(ln * rn)) continue;
                converged = false;
                const double zeta = (rn - ln) / (2.0 * cross);
                const double tangent = zeta == 0.0 ? 1.0 : copysign(1.0, zeta) / (fabs(zeta) + hypot(zeta, 1.0));
                const double cosine = 1.0 / sqrt(fma(tangent, tangent, 1.0));
                const double sine = cosine * tangent;
#pragma unroll
                for (int row = 0; row < COLUMNS; ++row) {
                    const double lv = mat[left * 5 + row], rv = mat[right * 5 + row];
                    mat[left * 5 + row] = fma(cosine, lv, -sine * rv);
                    mat[right * 5 + row] = fma(sine, lv, cosine * rv);
                }
            }
        }
        if (converged) break;
    }
    if (!converged) return -1;
#pragma unroll
    for (int c = 0; c < 5; ++c) sv[c] = 0.0;
#pragma unroll
    for (int c = 0; c < COLUMNS; ++c) {
        double sq = 0.0;
#pragma unroll
        for (int row = 0; row < COLUMNS; ++row) {
            const double v = mat[c * 5 + row];
            sq += v * v;
        }
        sv[c] = sqrt(sq);
    }
    // descending order (sorting network over the fixed width)
#pragma unroll
    for (int i = 0; i < COLUMNS; ++i)
#pragma unroll
        for (int j = 0; j + 1 < COLUMNS - i; ++j) {
            const double a = sv[j], b = sv[j + 1];
            sv[j] = a < b ? b : a;
            sv[j + 1] = a < b ? a : b;
        }
    return 0;
}

// Expands sum_k coef_k p_k(z) into monomials of z (p_0 = 1, p_1 = z - alpha_0,
// p_{k+1} = (z - alpha_k) p_k - beta_k p_{k-1}) and stores z's affine map.
__device__ __forceinline__ void set_evaluation_form(Model &m, int degree) {
    double previous[5] = {0, 0, 0, 0, 0}, current[5] = {1, 0, 0, 0, 0};
#pragma unroll
    for (int j = 0; j < 5; ++j) m.hz[j] = 0.0;
    m.hz[0] = m.coef[0];
#pragma unroll
    for (int k = 0; k < kMaxDegree; ++k) {
        if (k >= degree) break;
        double next[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int j = 0; j < 4; ++j) next[j + 1] += current[j];
#pragma unroll
        for (int j = 0; j < 5; ++j) next[j] = fma(-m.alpha[k], current[j], next[j]) - m.beta[k] * previous[j];
#pragma unroll
        for (int j = 0; j < 5; ++j) {
            m.hz[j] = fma(m.coef[k + 1], next[j], m.hz[j]);
            previous[j] = current[j];
            current[j] = next[j];
        }
    }
    m.degree = degree;
    const double inv_scale = degree == 0 ? 0.0 : 1.0 / m.scale;
    m.a = inv_scale / m.K;
    m.b = degree == 0 ? 0.0 : -m.mean * inv_scale;
}

// Rank gate and condition number of a completed fit (diagnose_design, regression.rs:540-577).
// Returns 0 and sets rank/condition when the fit is accepted, else a FOHIP_E_* code with its
// error detail.
__device__ __forceinline__ int gate_fit(Model &m, const FitParams &fp, int64_t &index, double &observed,
                                        double &limit) {
    double tr[25], sv[5];
    const int columns = fp.degree + 1;
    index = 0;
    observed = limit = 0.0;
    const int bad = orthonormal_transform(fp.degree, m, tr);
    if (bad) {
        index = bad - 1;
        return FOHIP_E_NON_FINITE_FIT;
    }
    int jr = 0;
    switch (columns) {
        case 1: jr = jacobi_singular_values<1>(tr, sv); break;
        case 2: jr = jacobi_singular_values<2>(tr, sv); break;
        case 3: jr = jacobi_singular_values<3>(tr, sv); break;
        case 4: jr = jacobi_singular_values<4>(tr, sv); break;
        default: jr = jacobi_singular_values<5>(tr, sv); break;
    }
    if (jr < 0) {
        index = columns;
        return FOHIP_E_JACOBI_NOT_CONVERGED;
    }
    if (jr > 0) {
        index = jr - 1;
        return FOHIP_E_NON_FINITE_FIT;
    }
    const double largest = sv[0], threshold = fp.rank_tol * largest;
    int rank = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) rank += (c < columns && sv[c] > threshold) ? 1 : 0;
    if (rank < columns) {
        index = rank;
        observed = columns;
        limit = fp.rank_tol;
        return FOHIP_E_RANK_DEFICIENT;
    }
    double smallest = sv[0];
#pragma unroll
    for (int c = 1; c < 5; ++c)
        if (c < columns) smallest = sv[c];
    const double condition = largest / smallest;
    if (!isfinite(condition)) {
        index = fp.degree;
        return FOHIP_E_NON_FINITE_FIT;
    }
    if (condition > fp.cond_limit) {
        index = fp.degree;
        observed = condition;
        limit = fp.cond_limit;
        return FOHIP_E_CONDITION_LIMIT;
    }
    m.rank = rank;
    m.condition = condition;
    return 0;
}

// Rank gate, condition number and acceptance of a completed fit. With DEFER the rule is accepted
// now and the gate runs later (k_fit_gate): the next node's fit uses only the coefficients, and a
// rule that fails the gate fails its row with the same error key.
template <bool DEFER>
__device__ void diagnose_and_accept(Model &m, const FitParams &fp, RowError *e, uint32_t step) {
    if (DEFER) {
        m.gate = kGatePending;
        set_evaluation_form(m, fp.degree);
        m.kind = kRuleRegression;
        return;
    }
    double tr[25], sv[5];
    const int columns = fp.degree + 1;
    const int bad = orthonormal_transform(fp.degree, m, tr);
    if (bad) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_NON_FINITE_FIT, bad - 1, 0.0, 0.0);
        return;
    }
    int jr = 0;
    switch (columns) {
        case 1: jr = jacobi_singular_values<1>(tr, sv); break;
        case 2: jr = jacobi_singular_values<2>(tr, sv); break;
        case 3: jr = jacobi_singular_values<3>(tr, sv); break;
        case 4: jr = jacobi_singular_values<4>(tr, sv); break;
        default: jr = jacobi_singular_values<5>(tr, sv); break;
    }
    if (jr < 0) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_JACOBI_NOT_CONVERGED, columns, 0.0, 0.0);
        return;
    }
    if (jr > 0) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_NON_FINITE_FIT, jr - 1, 0.0, 0.0);
        return;
    }
    const double largest = sv[0], threshold = fp.rank_tol * largest;
    int rank = 0;
#pragma unroll
    for (int c = 0; c < 5; ++c) rank += (c < columns && sv[c] > threshold) ? 1 : 0;
    if (rank < columns) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_RANK_DEFICIENT, rank, columns, fp.rank_tol);
        return;
    }
    double smallest = sv[0];
#pragma unroll
    for (int c = 1; c < 5; ++c)
        if (c < columns) smallest = sv[c];
    const double condition = largest / smallest;
    if (!isfinite(condition)) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_NON_FINITE_FIT, fp.degree, 0.0, 0.0);
        return;
    }
    if (condition > fp.cond_limit) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_CONDITION_LIMIT, fp.degree, condition, fp.cond_limit);
        return;
    }
    m.rank = rank;
    m.condition = condition;
    set_evaluation_form(m, fp.degree);
    m.kind = kRuleRegression;
}

// One Stieltjes step (regression.rs:448-514) from the reduced moments of p_k.
template <bool DEFER>
__device__ void stieltjes_step(Model &m, const FitParams &fp, RowError *e, uint32_t step, int k, double norm,
                               double projection, double weighted) {
    if (!isfinite(norm) || !isfinite(projection) || !isfinite(weighted)) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_NON_FINITE_FIT, k, 0.0, 0.0);
        return;
    }
    if (norm <= 0.0 || sqrt(norm / static_cast<double>(m.itm)) <= fp.rank_tol) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_RANK_DEFICIENT, k, fp.degree + 1, fp.rank_tol);
        return;
    }
    m.norm[k] = norm;
    m.coef[k] = projection / norm;
    if (!isfinite(m.coef[k])) {
        m.kind = kRuleNone;
        fail_row(e, step, FOHIP_E_NON_FINITE_FIT, k, 0.0, 0.0);
        return;
    }
    if (k < fp.degree) {
        m.alpha[k] = weighted / norm;
        m.beta[k] = k == 0 ? 0.0 : norm / m.prev_norm;
        if (!isfinite(m.alpha[k]) || !isfinite(m.beta[k])) {
            m.kind = kRuleNone;
            fail_row(e, step, FOHIP_E_NON_FINITE_FIT, k, 0.0, 0.0);
            return;
        }
        m.prev_norm = norm;
    } else {
        diagnose_and_accept<DEFER>(m, fp, e, step);
    }
}

__global__ __launch_bounds__(kBlock) void k_fit_reduce_moments(DevTables t, FitParams fp, uint32_t step, int k,
                                                               const uint32_t *__restrict__ rows,
                                                               double *__restrict__ partials,
                                                               Model *__restrict__ models,
                                                               RowError *__restrict__ errors) {
    __shared__ double scratch[3 * kWaves];
    const uint32_t srow = rows[blockIdx.x];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    if (step >= g.fit_rows) return;
    const uint32_t local = srow - g.strike_begin;
    const uint32_t r = g.obs - 2 - step;
    Model &m = models[g.model_base + local * g.obs + r];
    if (row_dead(errors, srow) || m.kind != kRulePending) return;
    double sums[3] = {0.0, 0.0, 0.0};
    for (uint32_t c = threadIdx.x; c < g.chunks_t; c += kBlock) {
        const double *in = partial_slot(partials, g, local, c);
        sums[0] += in[0];
        sums[1] += in[1];
        sums[2] += in[2];
    }
    block_sum<3>(sums, scratch);
    if (threadIdx.x == 0) stieltjes_step<false>(m, fp, &errors[srow], step, k, sums[0], sums[1], sums[2]);
}

// Fused backward step j over exercise row r = obs - 2 - j in ONE pass over HBM.
// Cashflows are stored discounted to the origin, v_o = cashflow * exp(logD(stop)), so a
// path's stored value changes only when it exercises; the response at row r is v_o / exp(logD(r)),
// the reference's step-by-step discounted value (policy.rs:1049-1063) in exact arithmetic.
// The pass applies row r + 1's rule (policy.rs:570-601) and accumulates the power sums
// P_j = sum (x - c)^j, j <= 2d, and G_j = sum y (x - c)^j, j <= d, of the moneyness x = S/K of
// in-the-money paths around a pivot c (row r + 1's fitted mean, else 1). The reduce kernel
// recentres them on the exact mean, which merges the Welford pass (regression.rs:279-297) and the
// Gram pass into one read of each value column.
// Storage (Store) is binary64 or binary32; moneyness, continuation, exercise decision, responses
// and power sums are binary64. Decisions made in binary32 arithmetic flip often enough to move
// prices by several tenths of a standard error; binary32 storage alone flips rare near-tie
// decisions, which backward induction then amplifies.
// Block = (2048-path chunk, set of up to 8 strikes): every thread keeps its 8 paths' spot values
// of rows r and r + 1 in registers for all strikes of the set; each wave reduces its sums through
// an LDS transpose, and one barrier at the end combines the waves in fixed order. The kernel is
// held to 128 VGPRs (4 waves per SIMD; 33.5 KB of LDS lets 4 blocks share a CU), so a C2 step's
// 4096 blocks run in 4 full rounds of the 1024 block slots.
__device__ __noinline__ void record_fit_errors(RowError *e, uint32_t step, uint32_t base, uint32_t n,
                                               uint32_t bad_bits, uint32_t nonfinite_bits);

template <typename Store, bool WIN>
__global__ __launch_bounds__(kBlock, 4) void k_fit_bulk_step(DevTables t, FitParams fp, uint32_t step,
                                                          const uint2 *__restrict__ items, uint32_t nitems,
                                                          uint32_t sets, const Store *__restrict__ paths,
                                                          Store *__restrict__ values, const Model *__restrict__ models,
                                                          double *__restrict__ partials,
                                                          RowError *__restrict__ errors) {
    __shared__ double red[kWaves][kBulkStrikes][16];
    __shared__ double transpose[kWaves][14 * 66];
    uint32_t wi, set;
    if (!xcd_item_column(nitems, sets, wi, set)) return;
    wi = serpentine_item(step, nitems, wi);
    const uint2 item = load_item(items, wi);
    const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + item.x);
    const uint32_t strike_begin = g->strike_begin, strike_end = g->strike_end;
    const uint32_t s_first = strike_begin + set * kBulkStrikes;
    if (s_first >= strike_end || step >= g->fit_rows) return;
    const int ns = static_cast<int>(min(static_cast<uint32_t>(kBulkStrikes), strike_end - s_first));
    const uint32_t obs = g->obs, r = obs - 2 - step, n = g->nt;
    const Store *S_r = paths + g->path_base + static_cast<uint64_t>(r) * n;
    const Store *S_next = S_r + n;
    const uint32_t base = item.y * kBulkChunk + threadIdx.x;
    Store s0v[kBulkPathsPerThread], s1v[kBulkPathsPerThread];
#pragma unroll
    for (int k = 0; k < kBulkPathsPerThread; ++k) {
        const uint32_t p = base + k * kBlock;
        s0v[k] = p < n ? S_r[p] : Store(0);
        // row r + 1 is read here for the last time (it was row r of the previous step): a streaming
        // load leaves the caches to row r and the cashflows, which the next step reads again
        s1v[k] = p < n ? __builtin_nontemporal_load(S_next + p) : Store(0);
    }
    const uint32_t row_base = g->row_base;
    const double d_next = *uniform(t.disc0 + row_base + r + 1);
    const double inv_d = 1.0 / *uniform(t.disc0 + row_base + r);
    const uint32_t model_row = g->model_base + r + 1;
    const uint64_t value_base = g->value_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // rows that already failed are skipped; the key is written by other blocks, so it is read
    // through the vector path once for the whole set
    uint32_t alive = 0;
    for (int i = 0; i < ns; ++i)
        alive |= (errors[s_first + i].key == kNoError ? 1u : 0u) << i;
    alive = __builtin_amdgcn_readfirstlane(alive);
    // cashflows of the next strike are loaded while the current one is processed; a chunk that
    // lies wholly inside the population (every chunk but a ragged last one) loads without masks
    Store vnext[kBulkPathsPerThread];
    const bool full_chunk = step > 0 && bulk_chunk_is_full(item.y, n);   // wave-uniform
    auto load_values = [&](int i, Store (&dst)[kBulkPathsPerThread]) {
        const Store *V = values + value_base + static_cast<uint64_t>(s_first + i - strike_begin) * n;
        if (full_chunk) {
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) dst[k] = V[base + k * kBlock];
        } else {
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) {
                const uint32_t p = base + k * kBlock;
                dst[k] = (step > 0 && p < n) ? V[p] : Store(0);
            }
        }
    };
    load_values(0, vnext);
    for (int i = 0; i < kBulkStrikes; ++i) {
        if (i >= ns) break;
        Store vv[kBulkPathsPerThread];
#pragma unroll
        for (int k = 0; k < kBulkPathsPerThread; ++k) vv[k] = vnext[k];
        if (i + 1 < ns) load_values(i + 1, vnext);
        const uint32_t srow = s_first + i, local = srow - strike_begin;
        double acc[14];
#pragma unroll
        for (int q = 0; q < 14; ++q) acc[q] = 0.0;
        const FOHIP_UNIFORM Model *mn = uniform(models + model_row + local * obs);
        const bool apply = step > 0 && mn->kind == kRuleRegression;   // wave-uniform
        const double pivot = apply ? mn->mean : 1.0;
        const double K = *uniform(t.strikes + srow);
        const double invK = *uniform(t.inv_strikes + srow);
        uint32_t tie_wave = 0;   // near ties of the wave at this strike, counted with its ballots
        if (alive & (1u << i)) {
            Store *V = values + value_base + static_cast<uint64_t>(local) * n;
            // Straight-line slots: lanes past n hold zero spots, which are never in the money and
            // never exercise, so only the stores need the bound.
            double v[kBulkPathsPerThread];
            bool bad[kBulkPathsPerThread];
            if (step == 0) {
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    const uint32_t p = base + k * kBlock;
                    const double s1 = s1v[k];
                    const Store vf = static_cast<Store>((s1 > K ? s1 - K : 0.0) * d_next);
                    if (p < n) V[p] = vf;
                    v[k] = vf;
                    bad[k] = false;
                }
            } else if (apply) {
                // continuation_fast's operations with its two leading addends held in vector
                // registers for the strike: a VOP3 FMA reads at most one scalar operand, so they
                // would otherwise be copied once per path slot
                double rule_b = mn->b, rule_h3 = mn->hz[3];
                asm volatile("" : "+v"(rule_b), "+v"(rule_h3));
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    const uint32_t p = base + k * kBlock;
                    const double s1 = s1v[k];
                    const double z = fma(s1, mn->a, rule_b);
                    double c = fma(mn->hz[4], z, rule_h3);
                    c = fma_uniform_addend(c, z, mn->hz[2]);
                    c = fma_uniform_addend(c, z, mn->hz[1]);
                    const double cont = fma_uniform_addend(c, z, mn->hz[0]);
                    const bool cand = s1 > K;
                    const bool finite = isfinite(cont);
                    const bool ex = cand && finite && should_exercise(fp.tie, s1 - K, cont);
                    // an exact tie implies a finite continuation, so its count is the AND of two
                    // compare masks
                    tie_wave += static_cast<uint32_t>(
                        WIN ? __popcll(__ballot(cand && finite && near_tie(s1 - K, cont, fp.tie_window)))
                            : __popcll(__ballot((s1 - K) == cont) & __ballot(cand)));
                    const Store vf = static_cast<Store>((s1 - K) * d_next);
                    if (ex) V[p] = vf;
                    v[k] = ex ? static_cast<double>(vf) : static_cast<double>(vv[k]);
                    bad[k] = cand && !finite;
                }
            } else {
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    v[k] = vv[k];
                    bad[k] = false;
                }
            }
            // failures are rare: one wave-uniform test for all slots
            bool any = false;
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) any = any | bad[k] | !isfinite(v[k]);   // no short circuit: no exec-mask branches
            if (__ballot(any) != 0ull) {
                uint32_t bad_bits = 0, nonfinite_bits = 0;
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    bad_bits |= (bad[k] ? 1u : 0u) << k;
                    nonfinite_bits |= (!isfinite(v[k]) ? 1u : 0u) << k;
                }
                record_fit_errors(&errors[srow], step, base, n, bad_bits, nonfinite_bits);
            }
            // pivoted power sums over in-the-money paths: an out-of-the-money lane adds exact zeros
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) {
                const double s0 = s0v[k];
                const bool itm = s0 > K;
                const double u = fma(s0, invK, -pivot);
                double pw = itm ? 1.0 : 0.0;
                // the response masked by a multiply (exact: v is a finite payoff >= 0 wherever the
                // row is still alive, and a non-finite v has already failed the row)
                const double y = (v[k] * inv_d) * pw;
#pragma unroll
                for (int j = 0; j <= 8; ++j) {
                    acc[j] += pw;
                    if (j <= 4) acc[9 + j] = fma(y, pw, acc[9 + j]);
                    pw *= u;
                }
            }
        }
        const double total = wave_sums_lds<14>(acc, transpose[wave]);
        if ((lane & 3) == 0 && lane < 56) red[wave][i][lane >> 2] = total;
        if (lane == 0) {
            red[wave][i][14] = pivot;
            red[wave][i][15] = static_cast<double>(tie_wave);
        }
    }
    __syncthreads();
    // slots 0..13 power sums, 14 the pivot, 15 the tied rule applications (an integer)
    for (int idx = threadIdx.x; idx < ns * 16; idx += kBlock) {
        const int i = idx / 16, q = idx % 16;
        const double v = q == 14 ? red[0][i][14] : (red[0][i][q] + red[1][i][q]) + (red[2][i][q] + red[3][i][q]);
        const uint32_t local = s_first + i - strike_begin;
        partials[static_cast<uint64_t>(g->part_base_t + local * g->chunks_t + item.y) * kPartialStride + q] = v;
    }
}

// The fused backward step with stopping-row cashflows. A path's cashflow for a strike is a function
// of its stopping row alone, v_o = max(S[stop] - K, 0) exp(logD(stop)), with the spot S[stop] of
// the group's own training matrix; so each (strike, path) keeps one byte, e = obs - 1 - stop
// (0: expiry, the state every path starts in; a path exercising in step j stops at row
// r + 1 = obs - 1 - j, i.e. e = j). The response of an in-the-money path at row r is that cashflow
// rebuilt from a gather of S[stop] (all strikes of a group share its spot matrix, and the strike
// sets of one path chunk run back to back on one XCD, so the gathers mostly hit its L2) over
// exp(logD(r)). The arithmetic is the value-column kernel's operation for operation, so the two
// encodings give bitwise the same power sums; the stopping bytes cost 1 B of HBM per strike and
// path and node instead of the binary64 column's 8 (plus 8 B per exercise rewrite).
template <typename Store, bool WIN>
__global__ __launch_bounds__(kBlock, 4) void k_fit_stop_step(DevTables t, FitParams fp, uint32_t step,
                                                          const uint2 *__restrict__ items, uint32_t nitems,
                                                          uint32_t sets, const Store *__restrict__ paths,
                                                          uint8_t *__restrict__ stops, const Model *__restrict__ models,
                                                          double *__restrict__ partials,
                                                          RowError *__restrict__ errors) {
    __shared__ double red[kWaves][kBulkStrikes][16];
    __shared__ double transpose[kWaves][14 * 66];
    __shared__ double disc_rows[256];   // exp(logD) of the group's observation rows (obs <= 255)
    uint32_t wi, set;
    if (!xcd_item_column(nitems, sets, wi, set)) return;
    wi = serpentine_item(step, nitems, wi);
    const uint2 item = load_item(items, wi);
    const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + item.x);
    const uint32_t strike_begin = g->strike_begin, strike_end = g->strike_end;
    const uint32_t s_first = strike_begin + set * kBulkStrikes;
    if (s_first >= strike_end || step >= g->fit_rows) return;
    const int ns = static_cast<int>(min(static_cast<uint32_t>(kBulkStrikes), strike_end - s_first));
    const uint32_t obs = g->obs, r = obs - 2 - step, n = g->nt;
    const Store *S_group = paths + g->path_base;
    const Store *S_r = S_group + static_cast<uint64_t>(r) * n;
    const Store *S_next = S_r + n;
    const uint32_t base = item.y * kBulkChunk + threadIdx.x;
    const uint32_t row_base = g->row_base;
    for (uint32_t i = threadIdx.x; i < obs; i += kBlock) disc_rows[i] = t.disc0[row_base + i];
    Store s0v[kBulkPathsPerThread], s1v[kBulkPathsPerThread];
#pragma unroll
    for (int k = 0; k < kBulkPathsPerThread; ++k) {
        const uint32_t p = base + k * kBlock;
        s0v[k] = p < n ? S_r[p] : Store(0);
        // row r + 1 is read here for the last time (it was row r of the previous step): a streaming
        // load leaves the caches to row r and the cashflows, which the next step reads again
        s1v[k] = p < n ? __builtin_nontemporal_load(S_next + p) : Store(0);
    }
    const double d_next = *uniform(t.disc0 + row_base + r + 1);
    const double inv_d = 1.0 / *uniform(t.disc0 + row_base + r);
    const uint32_t model_row = g->model_base + r + 1;
    const uint64_t stop_base = g->value_base;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t alive = 0;
    for (int i = 0; i < ns; ++i)
        alive |= (errors[s_first + i].key == kNoError ? 1u : 0u) << i;
    alive = __builtin_amdgcn_readfirstlane(alive);
    // stopping bytes of the next strike are loaded while the current one is processed (lanes past
    // n read nothing and keep e = 0, whose row they never gather)
    uint32_t enext[kBulkPathsPerThread];
    auto load_stops = [&](int i, uint32_t (&dst)[kBulkPathsPerThread]) {
        const uint8_t *E = stops + stop_base + static_cast<uint64_t>(s_first + i - strike_begin) * n;
#pragma unroll
        for (int k = 0; k < kBulkPathsPerThread; ++k) {
            const uint32_t p = base + k * kBlock;
            dst[k] = (step > 0 && p < n) ? static_cast<uint32_t>(E[p]) : 0u;
        }
    };
    load_stops(0, enext);
    __syncthreads();   // disc_rows
    for (int i = 0; i < kBulkStrikes; ++i) {
        if (i >= ns) break;
        uint32_t ev[kBulkPathsPerThread];
#pragma unroll
        for (int k = 0; k < kBulkPathsPerThread; ++k) ev[k] = enext[k];
        if (i + 1 < ns) load_stops(i + 1, enext);
        const uint32_t srow = s_first + i, local = srow - strike_begin;
        double acc[14];
#pragma unroll
        for (int q = 0; q < 14; ++q) acc[q] = 0.0;
        const FOHIP_UNIFORM Model *mn = uniform(models + model_row + local * obs);
        const bool apply = step > 0 && mn->kind == kRuleRegression;   // wave-uniform
        const double pivot = apply ? mn->mean : 1.0;
        const double K = *uniform(t.strikes + srow);
        const double invK = *uniform(t.inv_strikes + srow);
        uint32_t tie_wave = 0;
        if (alive & (1u << i)) {
            uint8_t *E = stops + stop_base + static_cast<uint64_t>(local) * n;
            // the current cashflow of every path in the money at row r: its stopping row's spot,
            // gathered before the decision so that the loads overlap the continuation
            Store sg[kBulkPathsPerThread];
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) {
                const uint32_t p = base + k * kBlock;
                const bool need = static_cast<double>(s0v[k]) > K;   // false for lanes past n
                sg[k] = need ? S_group[static_cast<uint64_t>(obs - 1 - ev[k]) * n + p] : Store(0);
            }
            double v[kBulkPathsPerThread];
            bool bad[kBulkPathsPerThread];
            double rule_b = apply ? mn->b : 0.0, rule_h3 = apply ? mn->hz[3] : 0.0;
            asm volatile("" : "+v"(rule_b), "+v"(rule_h3));
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) {
                const uint32_t p = base + k * kBlock;
                const double s1 = s1v[k];
                bool ex = false, cand = false, finite = true;
                if (apply) {
                    const double z = fma(s1, mn->a, rule_b);
                    double c = fma(mn->hz[4], z, rule_h3);
                    c = fma_uniform_addend(c, z, mn->hz[2]);
                    c = fma_uniform_addend(c, z, mn->hz[1]);
                    const double cont = fma_uniform_addend(c, z, mn->hz[0]);
                    cand = s1 > K;
                    finite = isfinite(cont);
                    ex = cand && finite && should_exercise(fp.tie, s1 - K, cont);
                    tie_wave += static_cast<uint32_t>(
                        WIN ? __popcll(__ballot(cand && finite && near_tie(s1 - K, cont, fp.tie_window)))
                            : __popcll(__ballot((s1 - K) == cont) & __ballot(cand)));
                    if (ex) E[p] = static_cast<uint8_t>(step);
                }
                const uint32_t stop = obs - 1 - ev[k];
                const double held = static_cast<double>(sg[k]);
                const double vf = (held > K ? held - K : 0.0) * disc_rows[stop];
                v[k] = ex ? (s1 - K) * d_next : vf;
                bad[k] = cand && !finite;
            }
            bool any = false;
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) any = any | bad[k] | !isfinite(v[k]);
            if (__ballot(any) != 0ull) {
                uint32_t bad_bits = 0, nonfinite_bits = 0;
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    bad_bits |= (bad[k] ? 1u : 0u) << k;
                    nonfinite_bits |= (!isfinite(v[k]) ? 1u : 0u) << k;
                }
                record_fit_errors(&errors[srow], step, base, n, bad_bits, nonfinite_bits);
            }
#pragma unroll
            for (int k = 0; k < kBulkPathsPerThread; ++k) {
                const double s0 = s0v[k];
                const bool itm = s0 > K;
                const double u = fma(s0, invK, -pivot);
                double pw = itm ? 1.0 : 0.0;
                const double y = (v[k] * inv_d) * pw;
#pragma unroll
                for (int j = 0; j <= 8; ++j) {
                    acc[j] += pw;
                    if (j <= 4) acc[9 + j] = fma(y, pw, acc[9 + j]);
                    pw *= u;
                }
            }
        }
        const double total = wave_sums_lds<14>(acc, transpose[wave]);
        if ((lane & 3) == 0 && lane < 56) red[wave][i][lane >> 2] = total;
        if (lane == 0) {
            red[wave][i][14] = pivot;
            red[wave][i][15] = static_cast<double>(tie_wave);
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < ns * 16; idx += kBlock) {
        const int i = idx / 16, q = idx % 16;
        const double v = q == 14 ? red[0][i][14] : (red[0][i][q] + red[1][i][q]) + (red[2][i][q] + red[3][i][q]);
        const uint32_t local = s_first + i - strike_begin;
        partials[static_cast<uint64_t>(g->part_base_t + local * g->chunks_t + item.y) * kPartialStride + q] = v;
    }
}

// Rare path of the fused fit pass (non-finite predictions or cashflows), kept out of line so that
// its atomics weigh neither on the register allocation nor on the code of the hot loop.
__device__ __noinline__ void record_fit_errors(RowError *e, uint32_t step, uint32_t base, uint32_t n,
                                               uint32_t bad_bits, uint32_t nonfinite_bits) {
    for (int k = 0; k < kBulkPathsPerThread; ++k) {
        const uint32_t p = base + k * kBlock;
        if ((bad_bits >> k) & 1u) record_error(e, error_key(step - 1, kStageApply, p));
        if (p < n && ((nonfinite_bits >> k) & 1u)) record_error(e, error_key(step, kStageDiscount, p));
    }
}

// Stieltjes recurrence from standardized power moments mu_j = sum z^j (j <= 2d) and projections
// nu_j = sum y z^j (j <= d): the monic orthogonal basis p_k is carried as monomial coefficients,
// norms and alpha use the Hankel moments, projections the right-hand side. This is the LDL^T
// factorization of the Gram matrix in orthogonal-polynomial form, so the fitted rule has exactly
// the representation (alpha, beta, coef) of the FP64 method.
__device__ void stieltjes_from_moments(Model &m, const FitParams &fp, RowError *e, uint32_t step,
                                       const double (&mu)[9], const double (&nu)[5]) {
    double a_prev[5] = {0, 0, 0, 0, 0}, a_cur[5] = {1, 0, 0, 0, 0};
#pragma unroll
    for (int k = 0; k <= kMaxDegree; ++k) {
        if (k > fp.degree) break;
        double norm = 0.0, weighted = 0.0, projection = 0.0;
#pragma unroll
        for (int i = 0; i <= k; ++i) {
            projection += a_cur[i] * nu[i];
#pragma unroll
            for (int j = 0; j <= k; ++j) {
                norm += a_cur[i] * a_cur[j] * mu[i + j];
                weighted += a_cur[i] * a_cur[j] * mu[i + j + 1 < 9 ? i + j + 1 : 8];
            }
        }
        stieltjes_step<true>(m, fp, e, step, k, norm, projection, weighted);
        if (m.kind != kRulePending) return;
        double a_next[5] = {0, 0, 0, 0, 0};
#pragma unroll
        for (int i = 0; i < kMaxDegree; ++i) {
            if (i <= k) {
                a_next[i + 1] += a_cur[i];
                a_next[i] -= m.alpha[k] * a_cur[i];
                a_next[i] -= m.beta[k] * a_prev[i];
            }
        }
#pragma unroll
        for (int i = 0; i < 5; ++i) {
            a_prev[i] = a_cur[i];
            a_cur[i] = a_next[i];
        }
    }
}

// Binomial coefficients C(j, k), j, k <= 8.
__device__ __forceinline__ double binom9(int j, int k) {
    constexpr double c[9][9] = {{1, 0, 0, 0, 0, 0, 0, 0, 0},       {1, 1, 0, 0, 0, 0, 0, 0, 0},
                                {1, 2, 1, 0, 0, 0, 0, 0, 0},       {1, 3, 3, 1, 0, 0, 0, 0, 0},
                                {1, 4, 6, 4, 1, 0, 0, 0, 0},       {1, 5, 10, 10, 5, 1, 0, 0, 0},
                                {1, 6, 15, 20, 15, 6, 1, 0, 0},    {1, 7, 21, 35, 35, 21, 7, 1, 0},
                                {1, 8, 28, 56, 70, 56, 28, 8, 1}};
    return c[j][k];
}

// Support decision, exact location/scale and standardized moments from the pivoted power sums
// (binomial recentring), then the Stieltjes solve above. One wave per strike row reduces the chunk
// partials of (strike row, exercise row obs - 2 - step) in fixed order (lane-strided chunk sums,
// then an LDS transpose) and its lane 0 fits the rule.
__device__ __forceinline__ void solve_node_from_sums(Model &m, const FitParams &fp, RowError *err, uint32_t step,
                                                     double strike, const double (&acc)[14], double pivot);

__global__ __launch_bounds__(64) void k_fit_bulk_reduce(DevTables t, FitParams fp, uint32_t step,
                                                        const uint32_t *__restrict__ rows,
                                                        double *__restrict__ partials, Model *__restrict__ models,
                                                        RowError *__restrict__ errors) {
    __shared__ double buf[14 * 66];
    const uint32_t srow = rows[blockIdx.x];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    if (step >= g.fit_rows || row_dead(errors, srow)) return;
    const int lane = threadIdx.x & 63;
    const uint32_t local = srow - g.strike_begin;
    const uint32_t r = g.obs - 2 - step;
    double acc[14];
#pragma unroll
    for (int i = 0; i < 14; ++i) acc[i] = 0.0;
    // four chunks' loads in flight per round trip; the sums still add chunks c, c + 64, ... in order
    double ties = 0.0;   // tied applications: integers, exact in any order
    uint32_t c = lane;
    for (; c + 3 * 64 < g.chunks_b; c += 4 * 64) {
        double in[4][16];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double *slot = partial_slot(partials, g, local, c + u * 64);
#pragma unroll
            for (int i = 0; i < 16; ++i) in[u][i] = slot[i];
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
#pragma unroll
            for (int i = 0; i < 14; ++i) acc[i] += in[u][i];
            ties += in[u][15];
        }
    }
    for (; c < g.chunks_b; c += 64) {
        const double *in = partial_slot(partials, g, local, c);
#pragma unroll
        for (int i = 0; i < 14; ++i) acc[i] += in[i];
        ties += in[15];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) ties += __shfl_xor(ties, off, 64);
    const double total = wave_sums_lds<14>(acc, buf);
    if ((lane & 3) == 0 && lane < 56) buf[lane >> 2] = total;   // totals of values 0..13
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    if (lane != 0) return;
#pragma unroll
    for (int i = 0; i < 14; ++i) acc[i] = buf[i];
    const double pivot = partial_slot(partials, g, local, 0)[14];
    // The rule is solved in registers and stored once: solving in place would turn every field
    // update into a global store and every re-read into a dependent global load.
    Model m = {};
    solve_node_from_sums(m, fp, &errors[srow], step, t.strikes[srow], acc, pivot);
    m.ties = ties;
    models[g.model_base + local * g.obs + r] = m;
}

// Support decision, binomial recentring of the pivoted power sums and moment-form Stieltjes solve
// of one (strike, row) of the fused method; `acc` holds the 9 + 5 power sums about `pivot`.
__device__ __forceinline__ void solve_node_from_sums(Model &m, const FitParams &fp, RowError *err, uint32_t step,
                                                     double strike, const double (&acc)[14], double pivot) {
    const double nd = acc[0];
    if (!decide_support(m, fp, err, step, static_cast<uint32_t>(nd), strike)) return;
    double mu[9], nu[5];
    double mean = 0.0, scale = 1.0;
    if (fp.degree == 0) {
#pragma unroll
        for (int j = 0; j < 9; ++j) mu[j] = j == 0 ? nd : 0.0;
#pragma unroll
        for (int j = 0; j < 5; ++j) nu[j] = j == 0 ? acc[9] : 0.0;
    } else {
        const double d = acc[1] / nd;
        mean = pivot + d;
        const double m2 = acc[2] - acc[1] * d;
        scale = sqrt(fmax(m2, 0.0) / nd);
        if (!isfinite(mean) || !isfinite(scale) || scale <= 0.0) {
            m.kind = kRuleNone;
            fail_row(err, step, FOHIP_E_DEGENERATE_PREDICTOR, fp.degree, scale, 0.0);
            return;
        }
        // sum (x - mean)^j = sum_k C(j, k) (-d)^(j - k) P_k, divided by scale^j
        double negd[9];
        negd[0] = 1.0;
#pragma unroll
        for (int j = 1; j < 9; ++j) negd[j] = negd[j - 1] * (-d);
        const double inv_scale = 1.0 / scale;
        double sp = 1.0;
#pragma unroll
        for (int j = 0; j < 9; ++j) {
            double q = 0.0, gq = 0.0;
#pragma unroll
            for (int k = 0; k <= j; ++k) {
                q += binom9(j, k) * negd[j - k] * acc[k];
                if (j < 5) gq += binom9(j, k) * negd[j - k] * acc[9 + k];
            }
            mu[j] = q * sp;
            if (j < 5) nu[j] = gq * sp;
            sp *= inv_scale;
        }
        mu[0] = nd;
        mu[1] = 0.0;   // exactly centred
    }
    begin_fit(m, mean, scale);
    stieltjes_from_moments(m, fp, err, step, mu, nu);
}

// ------------------------------------------------------------------------------- fit sweep

// Buffer resources: the fourth descriptor word of a raw (stride 0) 32-bit-format buffer, and the
// cache-policy bit of a streaming load.
constexpr int kBufferWord3 = 0x00020000;
constexpr int kBufferNonTemporal = 2;

// A double that every lane of the wave holds (read from LDS), as a wave-uniform scalar.
__device__ __forceinline__ double wave_uniform(double x) {
    const unsigned long long b = static_cast<unsigned long long>(__double_as_longlong(x));
    const uint32_t lo = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b));
    const uint32_t hi = __builtin_amdgcn_readfirstlane(static_cast<uint32_t>(b >> 32));
    return __longlong_as_double(static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Element `lane` of a segment of `bytes` bytes, or zero past its end, read with a streaming
// (non-temporal) buffer load: the resource's range check does the masking.
__device__ __forceinline__ double load_streaming(const double *seg, uint32_t bytes, uint32_t lane) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<double *>(seg), 0, static_cast<int>(bytes), kBufferWord3);
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(rs, lane * 8u, 0, kBufferNonTemporal));
}
__device__ __forceinline__ float load_streaming(const float *seg, uint32_t bytes, uint32_t lane) {
    const __amdgpu_buffer_rsrc_t rs =
        __builtin_amdgcn_make_buffer_rsrc(const_cast<float *>(seg), 0, static_cast<int>(bytes), kBufferWord3);
    return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rs, lane * 4u, 0, kBufferNonTemporal));
}

// Persistent backward induction of single-strike groups (fused methods, value-column cashflows).
// A group's fused-fit chunks are split over a team of blocks. A block keeps its chunks' spot values
// of rows r + 1 and r in registers and their cashflows in LDS from the first backward step to the
// last, so a step reads row r from HBM and nothing else (the per-node launches read rows r and
// r + 1 and the cashflow column and write exercised cashflows back at every step); row r - 1 is
// loaded while the team gathers. A block is four worker waves, whose threads hold paths exactly as
// the per-node kernel's threads do, and a control wave: each chunk's partial sums are the per-node
// kernel's, operation for operation (paths per thread, accumulation order, wave transpose, combine
// of the four waves). The control wave writes the block's partial records with device-scope
// stores (through to the coherence point, so no cache writeback is needed), waits for them to
// complete and arrives at the team's counter; once every member has arrived it reads the team's
// records with device-scope loads, merges them in k_fit_bulk_reduce's order and solves the rule
// itself, so every member holds the same rule and takes the same branches, and only the leader
// stores rules and failures. The control wave holds no path data, so the solve's registers and the rows' never
// add up. A failure in any chunk (recorded with the per-node keys) or in the solve ends the
// group's sweep at that step for every member, as the per-node reduce skips a dead row. The final
// cashflows go to the value column for k_fit_final.
// Termination: a block takes entries by ticket from a device-memory counter only while it runs,
// and a team's entries hold consecutive tickets, so at any time every team but the one whose
// tickets are being handed out has all its members running and finishes; blocks it frees take
// the remaining tickets of that team. Every team completes once the device can hold the largest
// team's blocks at once (the host checks this against the CU count), whatever else runs beside
// the sweep and however many of the grid's blocks are resident. A wait longer than
// plan.spin_limit polls (seconds) raises the abort word, which every waiting member sees, and
// all blocks return.
constexpr int kSweepWorkers = kWaves;               // worker waves: one chunk's 256 threads
constexpr int kSweepBlock = (kSweepWorkers + 1) * 64;

template <typename Store, bool WIN, int Q>
__global__ __launch_bounds__(kSweepBlock, 3) void k_fit_sweep(DevTables t, FitParams fp, SweepPlan plan,
                                                              const Store *__restrict__ paths,
                                                              Store *__restrict__ values,
                                                              Model *__restrict__ models,
                                                              RowError *__restrict__ errors) {
    constexpr int kFields = 17;   // record fields written: 14 sums, pivot, ties, failure
    __shared__ double red[kSweepWorkers][Q][kSweepStride];
    __shared__ double transpose[kSweepWorkers][14 * 66];   // workers' transposes; the merge's buffer
    __shared__ Store cash[Q][kBulkPathsPerThread][kBlock];   // the chunks' cashflows, lane-contiguous
    __shared__ Model rule;          // the rule the team fitted last: row r + 1 of the next step
    __shared__ RowError sink;       // failures of this member's own solve
    __shared__ uint32_t gathered;   // 0 the team has arrived, 2 abort
    __shared__ uint32_t verdict;    // 0 next step, 1 the group's sweep ends here
    __shared__ uint32_t ticket;     // the entry the block runs next (>= entry_count: none left)
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    uint32_t *abort_word = plan.sync + plan.teams;
    if (wave == 0) {
        // ---------------------------------------------------------------- control wave
        uint32_t *tickets = abort_word + 1;
        for (;;) {
            if (lane == 0)
                ticket = __hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u
                             ? plan.entry_count
                             : atomicAdd(tickets, 1u);
            __syncthreads();   // every wave knows the entry
            const uint32_t e = ticket;
            if (e >= plan.entry_count) return;
            const FOHIP_UNIFORM SweepEntry *job = uniform(plan.entries + e);
            const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + job->group);
            const uint32_t c0 = job->chunk_begin, nq = job->chunk_end - job->chunk_begin;
            const uint32_t obs = g->obs, fit_rows = g->fit_rows, cb = g->chunks_b;
            const uint32_t srow = g->strike_begin;
            const double K = *uniform(t.strikes + srow);
            double *team = plan.partials + job->part_base * kSweepStride;
            uint32_t *arrivals = plan.sync + job->team;
            for (uint32_t j = 0; j < fit_rows; ++j) {
                const uint32_t r = obs - 2 - j;
                __syncthreads();   // the workers' chunk sums are in `red`
                double *recs = team + static_cast<uint64_t>(j & 1u) * cb * kSweepStride;
                for (uint32_t idx = lane; idx < nq * kFields; idx += 64) {
                    const uint32_t q = idx / kFields, f = idx % kFields;
                    const double x = f == 14 ? red[0][q][14]
                                             : (red[0][q][f] + red[1][q][f]) + (red[2][q][f] + red[3][q][f]);
                    __hip_atomic_store(recs + static_cast<uint64_t>(c0 + q) * kSweepStride + f, x,
                                       __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                }
                // the records are device-scope stores (written through to the coherence point, no
                // cache writeback needed): the arrival waits only for them to complete
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                uint32_t status = 0;
                if (lane == 0) {
                    atomicAdd(arrivals, 1u);
                    const uint32_t target = job->members * (j + 1);
                    uint32_t polls = 0;
                    while (__hip_atomic_load(arrivals, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
                        if ((++polls & 63u) == 0u &&
                            (polls > plan.spin_limit ||
                             __hip_atomic_load(abort_word, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
                            atomicOr(abort_word, 1u);
                            status = 2;
                            break;
                        }
                        __builtin_amdgcn_s_sleep(1);
                    }
                }
                asm volatile("" ::: "memory");   // the team's records are read after every arrival
                status = __builtin_amdgcn_readfirstlane(status);
                if (lane == 0) gathered = status;
                if (status == 0u) {
                    // k_fit_bulk_reduce's merge: lane-strided chunk sums in chunk order (two chunks
                    // in flight), then the LDS transpose; ties are integers and failures flags,
                    // exact in any order
                    double acc[14];
#pragma unroll
                    for (int i = 0; i < 14; ++i) acc[i] = 0.0;
                    double ties = 0.0, fails = 0.0;
                    uint32_t c = lane;
                    for (; c + 64 < cb; c += 2 * 64) {
                        double in[2][kFields];
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
                            const double *slot = recs + static_cast<uint64_t>(c + u * 64) * kSweepStride;
#pragma unroll
                            for (int i = 0; i < kFields; ++i)
                                in[u][i] = __hip_atomic_load(slot + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                        }
#pragma unroll
                        for (int u = 0; u < 2; ++u) {
#pragma unroll
                            for (int i = 0; i < 14; ++i) acc[i] += in[u][i];
                            ties += in[u][15];
                            fails += in[u][16];
                        }
                    }
                    for (; c < cb; c += 64) {
                        const double *slot = recs + static_cast<uint64_t>(c) * kSweepStride;
                        double in[kFields];
#pragma unroll
                        for (int i = 0; i < kFields; ++i)
                            in[i] = __hip_atomic_load(slot + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
#pragma unroll
                        for (int i = 0; i < 14; ++i) acc[i] += in[i];
                        ties += in[15];
                        fails += in[16];
                    }
#pragma unroll
                    for (int off = 32; off > 0; off >>= 1) ties += __shfl_xor(ties, off, 64);
                    const bool dead = __ballot(fails != 0.0) != 0ull;
                    double *buf = transpose[0];   // the workers are between their transposes
                    const double total = wave_sums_lds<14>(acc, buf);
                    if ((lane & 3) == 0 && lane < 56) buf[lane >> 2] = total;
                    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
                    __builtin_amdgcn_wave_barrier();
                    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
                    if (lane == 0) {
                        if (dead) {
                            verdict = 1u;   // a chunk failed: the row is dead, as the per-node reduce finds it
                        } else {
#pragma unroll
                            for (int i = 0; i < 14; ++i) acc[i] = buf[i];
                            // chunk 0's pivot (every chunk holds it)
                            const double pivot = __hip_atomic_load(recs + 14, __ATOMIC_RELAXED,
                                                                   __HIP_MEMORY_SCOPE_AGENT);
                            sink.key = kNoError;
                            Model m = {};
                            solve_node_from_sums(m, fp, &sink, j, K, acc, pivot);
                            m.ties = ties;
                            if (job->leader != 0u) {
                                models[g->model_base + r] = m;
                                if (sink.key != kNoError) {
                                    RowError &out = errors[srow];
                                    out.key = sink.key;
                                    out.code = sink.code;
                                    out.index = sink.index;
                                    out.observed = sink.observed;
                                    out.limit = sink.limit;
                                }
                            }
                            rule = m;
                            verdict = sink.key != kNoError ? 1u : 0u;
                        }
                    }
                }
                __syncthreads();   // `gathered`, `verdict` and `rule` are set
                if (status != 0u) return;
                if (verdict != 0u) break;
            }
        }
        return;
    }
    // -------------------------------------------------------------------- worker waves
    const int worker = wave - 1;
    const uint32_t tw = threadIdx.x - 64;   // the thread's place in a chunk, as in the per-node kernel
    for (;;) {
        __syncthreads();   // the control wave has taken the entry
        const uint32_t e = ticket;
        if (e >= plan.entry_count) return;
        const FOHIP_UNIFORM SweepEntry *job = uniform(plan.entries + e);
        const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + job->group);
        const uint32_t c0 = job->chunk_begin;
        const uint32_t nq = job->chunk_end - c0;
        const uint32_t n = g->nt, obs = g->obs, fit_rows = g->fit_rows;
        const uint32_t srow = g->strike_begin, row_base = g->row_base;
        const double K = *uniform(t.strikes + srow);
        const double invK = *uniform(t.inv_strikes + srow);
        const Store *S = paths + g->path_base;
        Store s1[Q][kBulkPathsPerThread], s0[Q][kBulkPathsPerThread];
        // A chunk slot's row segment is read through a buffer resource that starts at the slot's
        // first path and ends at path n: one offset register (the thread's) serves every load, and
        // lanes past n read zero from the hardware's range check, the per-node kernel's zero spots.
        auto load_row = [&](uint32_t row, Store (&dst)[Q][kBulkPathsPerThread]) {
            const Store *R = S + static_cast<uint64_t>(row) * n;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (static_cast<uint32_t>(q) >= nq) break;
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    const uint32_t first = (c0 + q) * kBulkChunk + k * kBlock;
                    const uint32_t bytes = first < n ? (n - first) * static_cast<uint32_t>(sizeof(Store)) : 0u;
                    dst[q][k] = load_streaming(R + first, bytes, tw);
                }
            }
        };
        load_row(obs - 1, s1);
        load_row(obs - 2, s0);
        bool alive = true;
        for (uint32_t j = 0; j < fit_rows; ++j) {
            const uint32_t r = obs - 2 - j;
            const double d_next = *uniform(t.disc0 + row_base + r + 1);
            const double inv_d = 1.0 / *uniform(t.disc0 + row_base + r);
            const bool apply = j > 0 && __builtin_amdgcn_readfirstlane(rule.kind) == kRuleRegression;
            const double pivot = apply ? wave_uniform(rule.mean) : 1.0;
            double ra = 0.0, rb = 0.0, h0 = 0.0, h1 = 0.0, h2 = 0.0, h3 = 0.0, h4 = 0.0;
            if (apply) {
                ra = wave_uniform(rule.a);
                rb = wave_uniform(rule.b);
                h0 = wave_uniform(rule.hz[0]);
                h1 = wave_uniform(rule.hz[1]);
                h2 = wave_uniform(rule.hz[2]);
                h3 = wave_uniform(rule.hz[3]);
                h4 = wave_uniform(rule.hz[4]);
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (static_cast<uint32_t>(q) >= nq) break;
                const uint32_t base = (c0 + q) * kBulkChunk + tw;
                double acc[14];
#pragma unroll
                for (int i = 0; i < 14; ++i) acc[i] = 0.0;
                uint32_t tie_wave = 0, bad_bits = 0, nonfinite_bits = 0;
                // path by path: the cashflow after row r + 1's decision, its failure bits, then its
                // power-sum terms (the per-node kernel's operations, in its order for every sum)
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    const double x1 = s1[q][k];
                    double v;
                    bool bad = false;
                    if (j == 0) {
                        const Store vf = static_cast<Store>((x1 > K ? x1 - K : 0.0) * d_next);
                        cash[q][k][tw] = vf;
                        v = vf;
                    } else if (apply) {
                        // the rule's two leading addends in vector registers, as in the per-node kernel
                        double rule_b = rb, rule_h3 = h3;
                        asm volatile("" : "+v"(rule_b), "+v"(rule_h3));
                        const double z = fma(x1, ra, rule_b);
                        double c = fma(h4, z, rule_h3);
                        c = fma_uniform_addend(c, z, h2);
                        c = fma_uniform_addend(c, z, h1);
                        const double cont = fma_uniform_addend(c, z, h0);
                        const bool cand = x1 > K;
                        const bool finite = isfinite(cont);
                        const bool ex = cand && finite && should_exercise(fp.tie, x1 - K, cont);
                        tie_wave += static_cast<uint32_t>(
                            WIN ? __popcll(__ballot(cand && finite && near_tie(x1 - K, cont, fp.tie_window)))
                                : __popcll(__ballot((x1 - K) == cont) & __ballot(cand)));
                        const Store vf = static_cast<Store>((x1 - K) * d_next);
                        const Store held = cash[q][k][tw];
                        v = ex ? static_cast<double>(vf) : static_cast<double>(held);
                        if (ex) cash[q][k][tw] = vf;
                        bad = cand && !finite;
                    } else {
                        v = cash[q][k][tw];
                    }
                    bad_bits |= (bad ? 1u : 0u) << k;
                    nonfinite_bits |= (!isfinite(v) ? 1u : 0u) << k;
                    const double x0 = s0[q][k];
                    const bool itm = x0 > K;
                    const double u = fma(x0, invK, -pivot);
                    double pw = itm ? 1.0 : 0.0;
                    const double y = (v * inv_d) * pw;
#pragma unroll
                    for (int jj = 0; jj <= 8; ++jj) {
                        acc[jj] += pw;
                        if (jj <= 4) acc[9 + jj] = fma(y, pw, acc[9 + jj]);
                        pw *= u;
                    }
                }
                // failures are rare: one wave-uniform test for all slots
                const bool failed = __ballot((bad_bits | nonfinite_bits) != 0u) != 0ull;
                if (failed) record_fit_errors(&errors[srow], j, base, n, bad_bits, nonfinite_bits);
                const double total = wave_sums_lds<14>(acc, transpose[worker]);
                if ((lane & 3) == 0 && lane < 56) red[worker][q][lane >> 2] = total;
                if (lane == 0) {
                    red[worker][q][14] = pivot;
                    red[worker][q][15] = static_cast<double>(tie_wave);
                    red[worker][q][16] = failed ? 1.0 : 0.0;
                }
            }
            __syncthreads();   // chunk sums to the control wave
            // row r is the next step's row r + 1; row r - 1 loads while the team gathers
#pragma unroll
            for (int q = 0; q < Q; ++q)
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) s1[q][k] = s0[q][k];
            if (j + 1 < fit_rows) load_row(r - 1, s0);
            __syncthreads();   // the team's rule, or the end of the sweep
            if (gathered != 0u) return;
            if (verdict != 0u) {
                alive = false;
                break;
            }
        }
        if (alive) {   // final cashflows for k_fit_final
            Store *V = values + g->value_base;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                if (static_cast<uint32_t>(q) >= nq) break;
#pragma unroll
                for (int k = 0; k < kBulkPathsPerThread; ++k) {
                    const uint32_t p = (c0 + q) * kBulkChunk + k * kBlock + tw;
                    if (p < n) V[p] = cash[q][k][tw];
                }
            }
        }
    }
}

// After the last fitted row: apply it, discount to the origin (policy.rs:1065-1079) and sum the
// cashflows for the valuation-origin decision (policy.rs:267-281). Stored values are either
// discounted to the last fitted node (FP64 method, the reference's representation and arithmetic)
// or values already discounted to the origin (fused methods, evaluation-form decisions).
// STOPS: the fused cashflows are stopping bytes (see k_fit_stop_step), rebuilt here per path.
template <typename Real, bool FUSED, bool STOPS = false>
__global__ __launch_bounds__(kBlock) void k_fit_final(DevTables t, FitParams fp, const uint2 *__restrict__ items,
                                                      uint32_t nitems, uint32_t columns, const Real *__restrict__ paths,
                                                      const Real *__restrict__ values, const uint8_t *__restrict__ stops,
                                                      const Model *__restrict__ models,
                                                      double *__restrict__ partials, RowError *__restrict__ errors,
                                                      unsigned long long *__restrict__ ties) {
    constexpr bool kBulk = FUSED;
    __shared__ double scratch[kWaves];
    uint32_t wi, local;
    if (!xcd_item_column(nitems, columns, wi, local)) return;
    const uint2 item = items[wi];
    const DevGroup &g = t.groups[item.x];
    if (g.strike_begin + local >= g.strike_end) return;
    const uint32_t srow = g.strike_begin + local;
    if (row_dead(errors, srow)) return;
    const double K = t.strikes[srow];
    const uint32_t n = g.nt;
    const uint32_t last_row = g.origin_row;   // smallest positive-time exercise row
    const Real *S_last = paths + g.path_base + static_cast<uint64_t>(g.fit_rows ? last_row : g.obs - 1) * n;
    const Real *V = values + g.value_base + static_cast<uint64_t>(local) * n;
    const Model &m = models[g.model_base + local * g.obs + last_row];
    const bool apply = g.fit_rows > 0 && m.kind == kRuleRegression;
    const bool stored_at_origin = kBulk && g.fit_rows > 0;
    const double disc = stored_at_origin ? 1.0 : g.disc_final;
    const double d_last = stored_at_origin ? t.disc0[g.row_base + last_row] : 1.0;
    double sum[1] = {0.0};
    uint32_t tie_count = 0;
    // every slot's spot and cashflow are loaded before the first is used, so the loads are in
    // flight together (the sums below still run over the slots in order)
    Real s1v[kPathsPerThread], vv[kPathsPerThread];
#pragma unroll
    for (int k = 0; k < kPathsPerThread; ++k) {
        const uint32_t p = item.y * kChunk + k * kBlock + threadIdx.x;
        s1v[k] = p < n ? S_last[p] : Real(0);
        vv[k] = (!STOPS && g.fit_rows != 0 && p < n) ? V[p] : Real(0);
    }
#pragma unroll
    for (int k = 0; k < kPathsPerThread; ++k) {
        const uint32_t p = item.y * kChunk + k * kBlock + threadIdx.x;
        if (p >= n) break;
        double v;
        const double s1 = s1v[k];
        if (g.fit_rows == 0) {
            v = s1 > K ? s1 - K : 0.0;
        } else {
            if constexpr (STOPS) {
                const uint32_t stop = g.obs - 1 - static_cast<uint32_t>(stops[g.value_base + static_cast<uint64_t>(local) * n + p]);
                const double held = paths[g.path_base + static_cast<uint64_t>(stop) * n + p];
                v = (held > K ? held - K : 0.0) * t.disc0[g.row_base + stop];
            } else {
                v = vv[k];
            }
            if (apply && s1 > K) {
                const double cont = kBulk ? continuation_fast(&m, s1) : predict_ref(&m, fp.degree, s1 / K);
                if (!isfinite(cont)) {
                    record_error(&errors[srow], error_key(g.fit_rows - 1, kStageApply, p));
                } else {
                    tie_count += is_tie(s1 - K, cont, fp.tie_window) ? 1u : 0u;
                    if (should_exercise(fp.tie, s1 - K, cont))
                        v = STOPS ? (s1 - K) * d_last
                                  : (kBulk ? static_cast<double>(static_cast<Real>((s1 - K) * d_last)) : s1 - K);
                }
            }
        }
        v *= disc;
        if (!isfinite(v)) record_error(&errors[srow], error_key(g.fit_rows, kStageDiscount, p));
        sum[0] += v;
    }
    if (tie_count != 0u) add_ties(ties + srow, tie_count);
    block_sum<1>(sum, scratch);
    if (threadIdx.x == 0 && g.origin_row) partial_slot(partials, g, local, item.y)[0] = sum[0];
}

__global__ __launch_bounds__(kBlock) void k_fit_reduce_origin(DevTables t, FitParams fp,
                                                              const uint32_t *__restrict__ rows,
                                                              double *__restrict__ partials,
                                                              Model *__restrict__ models,
                                                              RowError *__restrict__ errors) {
    __shared__ double scratch[kWaves];
    const uint32_t srow = rows[blockIdx.x];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    if (!g.origin_row || row_dead(errors, srow)) return;
    const uint32_t local = srow - g.strike_begin;
    double sum[1] = {0.0};
    for (uint32_t c = threadIdx.x; c < g.chunks_t; c += kBlock) sum[0] += partial_slot(partials, g, local, c)[0];
    block_sum<1>(sum, scratch);
    if (threadIdx.x != 0) return;
    Model &m = models[g.model_base + local * g.obs];
    const double estimate = sum[0] / static_cast<double>(g.nt);
    const double K = t.strikes[srow];
    const double intrinsic0 = g.S0 - K > 0.0 ? g.S0 - K : 0.0;
    m.kind = kRuleOrigin;
    m.estimate = estimate;
    m.exercise = should_exercise(fp.tie, intrinsic0, estimate) ? 1 : 0;
}

// ------------------------------------------------------------------------------- evaluation

// Fused regenerate-and-evaluate (policy.rs:642-736). One evaluation path per thread, 256-path
// chunks, a tile of TS strikes per block: the thread regenerates its path from the Philox
// counters in binary64, applies every strike's frozen rule in row order and the block reduces
// each strike's discounted payoff moments. FAST selects the evaluation form of the rules (FP32
// bulk method); otherwise the reference's recurrence and divisions are used (FP64 method).
// Control flow over intervals, rows and strikes is wave-uniform (a failed lane keeps stepping
// with its result discarded), so rules and tables are read with scalar loads.
// With GREEKS, tangents of S w.r.t. (S0, sigma, r, q) ride along the transitions and each
// strike's stopping row is held fixed (pathwise estimator under the frozen policy).
// Rare path of the evaluation (a non-finite prediction or payoff), out of line.
__device__ __noinline__ void record_eval_errors(RowError *e, uint32_t row, uint32_t p, bool bad, bool payoff_bad) {
    if (bad) record_error(e, error_key(kEvalStepBase + row, kStageEvalPrediction, p));
    if (payoff_bad) record_error(e, error_key(kEvalStepBase + row, kStageEvalPayoff, p));
}

// Waves per SIMD asked of an instantiation. The fused Greek evaluation over 6- and 8-strike tiles
// of stored normals runs at 5 or more (at most 96 VGPRs) rather than 4: the extra wave hides more
// of the decision chain's latency. With all four tangents the 6-strike form, the default, fits
// without spills and the 8-strike form spills its payoff tangents; with the two of delta and vega
// both fit in 80 VGPRs (6 to 7 waves). Single-strike tiles that regenerate their paths run at 6
// (80 VGPRs, which their LDS allows) when they carry at most two tangents; with all four they
// would spill and keep the compiler's own choice, as every other form does.
template <bool FAST, int TS, int GREEKS, bool STORED>
constexpr int eval_waves() {
    return (FAST && GREEKS && STORED && (TS == 6 || TS == 8)) ? 5 : ((TS == 1 && !STORED && GREEKS != 4) ? 6 : 1);
}

// Decision data of one (row, strike) as staged in LDS: K, a, b, hz0..hz4.
constexpr int kDecDoubles = 8;

// The fast continuation from a staged entry: the same operations, in the same order, as
// continuation_fast on the rule itself.
__device__ __forceinline__ double continuation_staged(const double *d, double spot) {
    const double z = fma(spot, d[1], d[2]);
    double c = fma(d[7], z, d[6]);
    c = fma(c, z, d[5]);
    c = fma(c, z, d[4]);
    return fma(c, z, d[3]);
}

// DEC: the tile's decision data (kDecDoubles per (row, strike), [row][strike]) and each row's rule
// kinds (4 bits per strike) are staged in LDS after the exercise histogram, so that a decision
// reads LDS broadcasts instead of waiting on scalar loads of rules that lie obs rules apart.
template <bool FAST, int TS, int GREEKS, bool ESCROW, bool STORED, bool DEC, bool WIN>
__global__ __launch_bounds__(kBlock, (eval_waves<FAST, TS, GREEKS, STORED>())) void k_evaluate(DevTables t, FitParams fp, const uint2 *__restrict__ items,
                                                     uint32_t nitems, uint32_t tiles, const Model *__restrict__ models,
                                                     double *__restrict__ partials, uint32_t *__restrict__ counts,
                                                     uint32_t *__restrict__ nosupport, RowError *__restrict__ errors,
                                                     unsigned long long *__restrict__ sim_fail, int use_lds_hist,
                                                     const double *__restrict__ znorm) {
    using Tan = typename std::conditional<FAST, float, double>::type;
    extern __shared__ __align__(16) uint32_t hist[];   // [TS][obs] when use_lds_hist, then DEC's tables
    __shared__ double sx[257];
    __shared__ float4 ssq[256];
    __shared__ ExpEntry se[32];
    constexpr int NACC = GREEKS ? 10 : 2;
    __shared__ double scratch[NACC * kWaves];
    __shared__ uint32_t tie_lds[TS];   // tied decisions of this tile, written with its partials
    uint32_t wi, tile;
    if (!xcd_item_column(nitems, tiles, wi, tile)) return;
    const uint2 item = load_item(items, wi);
    const FOHIP_UNIFORM DevGroup *g = uniform(t.groups + item.x);
    const uint32_t strike_begin = g->strike_begin, strike_end = g->strike_end;
    const uint32_t s_first = strike_begin + tile * TS;
    if (s_first >= strike_end) return;
    const int ns = static_cast<int>(min(static_cast<uint32_t>(TS), strike_end - s_first));
    const uint32_t obs = g->obs, last = obs - 1, intervals = g->intervals, n = g->ne;
    const uint32_t step_base = g->step_base, row_base = g->row_base;
    const uint32_t origin_row = g->origin_row, count_base = g->count_base;
    const int policy = g->policy;
    const double sigma = g->sigma;
    const uint32_t local0 = s_first - strike_begin;
    const Model *tile_models = models + g->model_base + local0 * obs;   // [strike][row]
    const int degree = fp.degree, tie = fp.tie;
    if (!STORED) load_zig(sx, ssq);
    for (int i = threadIdx.x; i < 32; i += kBlock) se[i] = g_exp_table[i];
    if (threadIdx.x < TS) tie_lds[threadIdx.x] = 0;
    __syncthreads();
    if (use_lds_hist) {
        for (uint32_t i = threadIdx.x; i < TS * obs; i += kBlock) hist[i] = 0;
        __syncthreads();
    }
    double *dec = nullptr;
    uint32_t *dec_kind = nullptr;
    if constexpr (DEC) {
        static_assert(TS <= 8, "rule kinds are packed 4 bits per strike into one word per row");
        dec = reinterpret_cast<double *>(hist + ((TS * obs + 3u) & ~3u));
        dec_kind = reinterpret_cast<uint32_t *>(dec + static_cast<size_t>(obs) * TS * kDecDoubles);
        for (uint32_t e = threadIdx.x; e < static_cast<uint32_t>(ns) * obs; e += kBlock) {
            const uint32_t i = e / obs, r = e - i * obs;
            const Model *m = tile_models + i * obs + r;
            double *d = dec + (static_cast<size_t>(r) * TS + i) * kDecDoubles;
            d[0] = m->K;
            d[1] = m->a;
            d[2] = m->b;
#pragma unroll
            for (int k = 0; k <= kMaxDegree; ++k) d[3 + k] = m->hz[k];
        }
        for (uint32_t r = threadIdx.x; r < obs; r += kBlock) {
            uint32_t word = 0;
            for (int i = 0; i < ns; ++i) word |= (static_cast<uint32_t>(tile_models[i * obs + r].kind) & 15u) << (4 * i);
            dec_kind[r] = word;
        }
        __syncthreads();
    }
    uint32_t live = 0;
#pragma unroll
    for (int i = 0; i < TS; ++i) {
        if (i < ns) {
            const FOHIP_UNIFORM Model *m0 = uniform(tile_models + i * obs);
            const bool origin_exercise = origin_row && m0->kind == kRuleOrigin && m0->exercise;
            const int healthy = __builtin_amdgcn_readfirstlane(errors[s_first + i].key == kNoError ? 1 : 0);
            if (!origin_exercise && healthy) live |= 1u << i;
        }
    }
    const uint32_t p = item.y * kEvalChunk + threadIdx.x;
    const int lane = threadIdx.x & 63;
    bool ok = p < n;
    uint32_t active = ok ? live : 0u;
    double pay[TS];
    Tan dpay[TS][GREEKS ? GREEKS : 1];
    // DEC: each lane keeps its stopping row per strike, 8 bits each (0xFF: no stop; the launch
    // guarantees obs <= 255), and the exercise histogram is added once at the end rather than
    // with a ballot and an atomic at every stop
    uint32_t stop_rows[(TS + 3) / 4];
#pragma unroll
    for (int h = 0; h < (TS + 3) / 4; ++h) stop_rows[h] = 0xFFFFFFFFu;
#pragma unroll
    for (int i = 0; i < TS; ++i) {
        pay[i] = 0.0;
#pragma unroll
        for (int a = 0; a < (GREEKS ? GREEKS : 1); ++a) dpay[i][a] = Tan(0);
    }
    __shared__ uint64_t ring[STORED ? 1 : kRingSlots * kBlock];
    RingStream s;
    if (!STORED) s.init(g->eval_seed, g->eval_domain, p, ring, threadIdx.x, kBlock);
    // STORED: this path's normals, time-major, written once by k_eval_normals (NaN marks an
    // exhausted draw budget)
    const double *zp = znorm + g->znorm_base + p;
    double spot = g->start;   // S0, or the escrowed prepaid forward F0
    double tan[4] = {1.0, 0.0, 0.0, 0.0};
    uint32_t row = origin_row;   // the origin row holds no regression rule
    for (uint32_t iv = 0; iv < intervals; ++iv) {
        double z = 0.0;
        if (STORED) {
            if (p < n) z = zp[static_cast<uint64_t>(iv) * n];
            if (!(z == z)) ok = false;
        } else {
            uint32_t draws = 0;
            s.top_up(ok);
            if (ok && !zignor_normal(s, sx, ssq, g_zig_f, FOHIP_ZIG_R, z, draws, se)) ok = false;
        }
        const FOHIP_UNIFORM StepRec *sr = uniform(t.steps + step_base + iv);
        double growth;
        const double next = gbm_advance(spot, sr->diffusion, sr->log_drift, z, growth, se);
        if (!isfinite(next)) ok = false;
        if (!ok) active = 0u;
        const double before = spot;
        spot = ok ? next : spot;
        if (GREEKS) {
            // an absorbed spot model path (S = 0) is never in the money again (K > 0), so its
            // tangents are never read and need no clearing; an escrowed path at F = 0 still
            // observes the prepaid cash and must carry zero tangents
            if (ESCROW && before == 0.0) {
                tan[0] = tan[1] = tan[2] = tan[3] = 0.0;
            } else {
                const double dt = sr->dt;
                tan[0] = tan[0] * growth;
                tan[1] = fma(tan[1], growth, spot * fma(sr->sqrt_dt, z, -sigma * dt));
                if (GREEKS == 4) {
                    tan[2] = fma(tan[2], growth, spot * dt);
                    tan[3] = fma(tan[3], growth, -spot * dt);
                }
            }
        }
        const uint32_t flags = sr->flags;
        const bool cash = flags & kNodeCash;
        const bool terminal = flags & kNodeTerminal;
        const double amount = sr->cash;   // 0 where the node pays none
        if (cash && terminal) {
            if (GREEKS && policy == 0 && !(spot > amount)) tan[0] = tan[1] = tan[2] = tan[3] = 0.0;
            spot = apply_cash<double>(policy, amount, spot);
        }
        if ((flags & kNodeObserved) && row < last) {
            const double disc = sr->disc0;   // this row's (the record of an observed node holds it)
            const double t_row = GREEKS ? sr->row_time : 0.0;
            // the exercise decision observes the stock: F + A(node) for escrowed groups
            const double S = ESCROW ? spot + *uniform(t.row_escrow + row_base + row) : spot;
            // discounted observed tangents, shared by every strike that stops on this row: with
            // several strikes per tile formed once per row (and kept from being recomputed per
            // strike), with one strike only inside the branch a stopping lane takes
            Tan od0 = Tan(0), od1 = Tan(0), od3 = Tan(0);
            double od2 = 0.0;
            auto form_od = [&]() {
                double o[4] = {tan[0], tan[1], tan[2], tan[3]};
                if (ESCROW) observed_tangents(tan, uniform(g->chain), *uniform(t.row_escrow_dr + row_base + row), o);
                od0 = static_cast<Tan>(o[0] * disc);
                od1 = static_cast<Tan>(o[1] * disc);
                if (GREEKS == 4) {
                    od2 = o[2] * disc;
                    od3 = static_cast<Tan>(o[3] * disc);
                }
            };
            if constexpr (GREEKS != 0 && TS > 1) {
                form_od();
                if constexpr (GREEKS == 4) asm volatile("" : "+v"(od0), "+v"(od1), "+v"(od3), "+v"(od2));
                else asm volatile("" : "+v"(od0), "+v"(od1));
            }
            const uint32_t kinds_row = DEC ? __builtin_amdgcn_readfirstlane(dec_kind[row]) : 0u;
            const double *dec_row = DEC ? dec + static_cast<size_t>(row) * TS * kDecDoubles : nullptr;
#pragma unroll
            for (int i = 0; i < TS; ++i) {
                if (i >= ns) break;
                const FOHIP_UNIFORM Model *m = DEC ? nullptr : uniform(tile_models + i * obs + row);
                const double *d = DEC ? dec_row + i * kDecDoubles : nullptr;
                const int kind = DEC ? static_cast<int>((kinds_row >> (4 * i)) & 15u) : m->kind;   // wave-uniform
                const double K = DEC ? d[0] : m->K;   // set on every fitted row
                const bool cand = ((active >> i) & 1u) && S > K;
                // branch-free: the continuation is formed for every row and used only where a
                // regression rule exists
                const bool reg = kind == kRuleRegression;
                const double intr = S - K;
                const double cont = DEC ? continuation_staged(d, S)
                                        : (FAST ? continuation_fast(m, S) : predict_ref(m, degree, S / K));
                const bool finite = isfinite(cont);
                const bool ex = reg && cand && finite && should_exercise(tie, intr, cont);
                const bool bad = reg && cand && !finite;
                const bool tied = reg && cand && finite && (WIN ? near_tie(intr, cont, fp.tie_window) : intr == cont);
                if constexpr (DEC) {
                    // state changes only where a lane stops or ties: a per-lane branch, which the
                    // wave skips when none does (its histogram entry is the lane's stopping row)
                    if (ex || bad || tied) {
                        if (tied) atomicAdd(&tie_lds[i], 1u);
                        const double pv = intr * disc;
                        if (ex) {
                            pay[i] = pv;
                            if constexpr (GREEKS != 0) {
                                if constexpr (TS == 1) form_od();
                                dpay[i][0] = od0;
                                dpay[i][1] = od1;
                                if constexpr (GREEKS == 4) {
                                    dpay[i][2] = static_cast<Tan>(od2 - t_row * pv);
                                    dpay[i][3] = od3;
                                }
                            }
                            const int sh = 8 * (i & 3);
                            stop_rows[i >> 2] = (stop_rows[i >> 2] & ~(0xFFu << sh)) | (row << sh);
                        }
                        if (ex || bad) active &= ~(1u << i);
                        const bool payoff_bad = ex && !isfinite(pv);
                        if (bad || payoff_bad) record_eval_errors(&errors[s_first + i], row, p, bad, payoff_bad);
                    }
                } else if (__ballot(ex || bad || tied) != 0ull) {
                    // state changes only where a lane stops (or a tie is counted): one wave-uniform branch
                    const unsigned long long tm = __ballot(tied);
                    if (tm != 0ull && lane == 0) atomicAdd(&tie_lds[i], static_cast<uint32_t>(__popcll(tm)));
                    const double pv = intr * disc;
                    pay[i] = ex ? pv : pay[i];
                    if constexpr (GREEKS != 0) {
                        if constexpr (TS == 1) form_od();
                        dpay[i][0] = ex ? od0 : dpay[i][0];
                        dpay[i][1] = ex ? od1 : dpay[i][1];
                        if constexpr (GREEKS == 4) {
                            dpay[i][2] = ex ? static_cast<Tan>(od2 - t_row * pv) : dpay[i][2];
                            dpay[i][3] = ex ? od3 : dpay[i][3];
                        }
                    }
                    active &= ~(((ex || bad) ? 1u : 0u) << i);
                    const bool payoff_bad = ex && !isfinite(pv);
                    if (__ballot(bad || payoff_bad) != 0ull)
                        record_eval_errors(&errors[s_first + i], row, p, bad, payoff_bad);
                    const unsigned long long exm = __ballot(ex);
                    if (exm != 0ull && lane == 0) {
                        const uint32_t c = static_cast<uint32_t>(__popcll(exm));
                        if (use_lds_hist)
                            atomicAdd(&hist[i * obs + row], c);
                        else
                            atomicAdd(&counts[count_base + (local0 + i) * obs + row], c);
                    }
                }
                if (kind == kRuleNoSupport) {   // wave-uniform
                    const unsigned long long cm = __ballot(cand);
                    if (cm != 0ull && lane == 0)
                        atomicAdd(&nosupport[count_base + (local0 + i) * obs + row], static_cast<uint32_t>(__popcll(cm)));
                }
            }
        }
        if (flags & kNodeObserved) row += 1;
        if (cash && !terminal) {
            if (GREEKS && policy == 0 && !(spot > amount)) tan[0] = tan[1] = tan[2] = tan[3] = 0.0;
            spot = apply_cash<double>(policy, amount, spot);
        }
    }
    if (p < n && !ok) atomicMin(sim_fail + item.x, static_cast<unsigned long long>(p));
    {
        const double disc = *uniform(t.disc0 + row_base + last);
        const double t_last = *uniform(t.row_time + row_base + last);
        double o[4] = {tan[0], tan[1], tan[2], tan[3]};   // expiry observes F_T itself (no prepaid cash)
        if (GREEKS && ESCROW) observed_tangents(tan, uniform(g->chain), 0.0, o);
#pragma unroll
        for (int i = 0; i < TS; ++i) {
            if (i >= ns) break;
            const double K = *uniform(t.strikes + s_first + i);
            const bool hold = (active >> i) & 1u;
            const double pv = (spot > K ? spot - K : 0.0) * disc;
            pay[i] = hold ? pv : pay[i];
            if (hold && !isfinite(pv))
                record_error(&errors[s_first + i], error_key(kEvalStepBase + last, kStageEvalPayoff, p));
            if constexpr (DEC) {
                const uint32_t sr = hold ? last : ((stop_rows[i >> 2] >> (8 * (i & 3))) & 0xFFu);
                if (sr != 0xFFu) {
                    if (use_lds_hist)
                        atomicAdd(&hist[i * obs + sr], 1u);
                    else
                        atomicAdd(&counts[count_base + (local0 + i) * obs + sr], 1u);
                }
            } else {
                const unsigned long long hm = __ballot(hold);
                if (hm != 0ull && lane == 0) {
                    const uint32_t c = static_cast<uint32_t>(__popcll(hm));
                    if (use_lds_hist)
                        atomicAdd(&hist[i * obs + last], c);
                    else
                        atomicAdd(&counts[count_base + (local0 + i) * obs + last], c);
                }
            }
            if constexpr (GREEKS != 0) {
                const bool itm = hold && spot > K;
                dpay[i][0] = itm ? static_cast<Tan>(o[0] * disc) : dpay[i][0];
                dpay[i][1] = itm ? static_cast<Tan>(o[1] * disc) : dpay[i][1];
                if constexpr (GREEKS == 4) {
                    dpay[i][2] = itm ? static_cast<Tan>(o[2] * disc - t_last * pv) : dpay[i][2];
                    dpay[i][3] = itm ? static_cast<Tan>(o[3] * disc) : dpay[i][3];
                }
            }
        }
    }
    // per-strike block reduction of (sum v, sum v^2[, sum dv_a, sum dv_a^2])
    const bool contributes_any = p < n && ok;
#pragma unroll
    for (int i = 0; i < TS; ++i) {
        if (i >= ns) break;
        const bool contributes = contributes_any && (live & (1u << i));
        double vals[NACC];
        const double v = contributes ? pay[i] : 0.0;
        vals[0] = v;
        vals[1] = v * v;
        if (GREEKS) {
#pragma unroll
            for (int a = 0; a < 4; ++a) {   // tangents not carried contribute zeros
                const double d = (a < GREEKS && contributes) ? static_cast<double>(dpay[i][a < GREEKS ? a : 0]) : 0.0;
                vals[2 + a] = d;
                vals[6 + a] = d * d;
            }
        }
        block_sum<NACC>(vals, scratch);
        if (threadIdx.x == 0) {
            double *out = partials + static_cast<uint64_t>(g->part_base_e + (local0 + i) * g->chunks_e + item.y) *
                                         kPartialStride;
#pragma unroll
            for (int a = 0; a < NACC; ++a) out[a] = vals[a];
            out[kPartialStride - 1] = static_cast<double>(tie_lds[i]);   // block_sum's barriers published it
        }
    }
    if (use_lds_hist) {
        __syncthreads();
        for (uint32_t idx = threadIdx.x; idx < static_cast<uint32_t>(ns) * obs; idx += kBlock) {
            const uint32_t c = hist[idx];
            if (c) atomicAdd(&counts[count_base + (local0 + idx / obs) * obs + idx % obs], c);
        }
    }
}

// Per strike: exact per-chunk (count, mean, M2) from the chunk sums, merged in chunk order
// (policy.rs:891-926 computes the same population moments sequentially).
__global__ __launch_bounds__(kBlock) void k_eval_reduce(DevTables t, const uint32_t *__restrict__ rows,
                                                        const double *__restrict__ partials, int greeks,
                                                        double *__restrict__ out) {
    __shared__ double scratch[9 * kWaves];
    __shared__ double mscratch[3 * kWaves];
    const uint32_t srow = rows[blockIdx.x];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    const uint32_t local = srow - g.strike_begin;
    Moments acc = {0.0, 0.0, 0.0};
    double gs[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};   // Greek sums, then the tied decisions
    for (uint32_t c = threadIdx.x; c < g.chunks_e; c += kBlock) {
        const double *in = partials + static_cast<uint64_t>(g.part_base_e + local * g.chunks_e + c) * kPartialStride;
        const double nc = static_cast<double>(min(static_cast<uint32_t>(kEvalChunk), g.ne - c * kEvalChunk));
        const double mean = in[0] / nc;
        const double m2 = fmax(in[1] - in[0] * mean, 0.0);
        acc = chan_merge(acc, Moments{nc, mean, m2});
        if (greeks)
#pragma unroll
            for (int a = 0; a < 8; ++a) gs[a] += in[2 + a];
        gs[8] += in[kPartialStride - 1];
    }
    acc = block_chan(acc, mscratch);
    block_sum<9>(gs, scratch);
    if (threadIdx.x == 0) {
        double *o = out + static_cast<uint64_t>(srow) * 16;
        const double nn = static_cast<double>(g.ne);
        o[0] = acc.mean;
        o[1] = acc.m2;
        o[2] = nn;
        for (int a = 0; a < 4; ++a) {
            const double mean = gs[a] / nn;
            const double var = fmax((gs[4 + a] - gs[a] * mean) / (nn - 1.0), 0.0);
            o[3 + a] = mean;
            o[7 + a] = sqrt(var / nn);
        }
        o[15] = gs[8];   // tied evaluation decisions (an integer)
    }
}

// Per strike row: the 12-column summary of try_from_parts (batch/pricing.rs:1130-1230) computed
// from the fitted rules and stopping counts, plus the first evaluation row whose declared
// missing-support rule met an active in-the-money path (policy.rs:684-703). One wave per strike
// row, its lanes striding over the rows: every column is a count, a sum of integers (the tie
// counts), a maximum, a minimum or a first index, so the result does not depend on the order.
__global__ __launch_bounds__(kBlock) void k_row_summary(DevTables t, const Model *__restrict__ models,
                                                        const uint32_t *__restrict__ counts,
                                                        const uint32_t *__restrict__ nosupport, uint32_t rows,
                                                        RowSummary *__restrict__ out) {
    const uint32_t srow = blockIdx.x * kWaves + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (srow >= rows) return;   // wave-uniform
    const DevGroup &g = t.groups[t.strike_group[srow]];
    const uint32_t local = srow - g.strike_begin;
    const uint32_t model0 = g.model_base + local * g.obs, count0 = g.count_base + local * g.obs;
    double ties = 0.0, max_condition = 0.0, estimate = 0.0;
    uint32_t regression = 0, no_support = 0, insufficient = 0, min_itm = 0xFFFFFFFFu, early = 0;
    uint32_t origin = 0, exercised = 0, first_nosupport = 0xFFFFFFFFu;
    for (uint32_t r = lane; r + 1 < g.obs; r += 64) {
        const Model &m = models[model0 + r];
        ties += m.ties;
        switch (m.kind) {
            case kRuleOrigin:
                origin = 1;
                exercised = m.exercise ? 1 : 0;
                estimate = m.estimate;
                break;
            case kRuleNoSupport:
                no_support += 1;
                break;
            case kRuleContinue:
                insufficient += 1;
                break;
            case kRuleRegression:
                regression += 1;
                min_itm = min(min_itm, m.itm);
                max_condition = fmax(max_condition, m.condition);
                break;
            default:
                break;
        }
        early += counts[count0 + r];
        if (first_nosupport == 0xFFFFFFFFu && nosupport[count0 + r] > 0) first_nosupport = r;
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        ties += __shfl_xor(ties, off, 64);
        max_condition = fmax(max_condition, __shfl_xor(max_condition, off, 64));
        regression += __shfl_xor(regression, off, 64);
        no_support += __shfl_xor(no_support, off, 64);
        insufficient += __shfl_xor(insufficient, off, 64);
        min_itm = min(min_itm, static_cast<uint32_t>(__shfl_xor(min_itm, off, 64)));
        early += __shfl_xor(early, off, 64);
        first_nosupport = min(first_nosupport, static_cast<uint32_t>(__shfl_xor(first_nosupport, off, 64)));
    }
    // the origin rule (row 0) lies in one lane
    const unsigned long long holder = __ballot(origin != 0u);
    if (holder != 0ull) {
        const int from = __ffsll(static_cast<long long>(holder)) - 1;
        exercised = __shfl(exercised, from, 64);
        estimate = __shfl(estimate, from, 64);
    }
    if (lane != 0) return;
    RowSummary rs;
    rs.origin_fitted = holder != 0ull ? 1 : 0;
    rs.origin_exercised = holder != 0ull && exercised ? 1 : 0;
    rs.origin_estimate = holder != 0ull ? estimate : 0.0;
    rs.regression_nodes = regression;
    rs.no_support_nodes = no_support;
    rs.insufficient_nodes = insufficient;
    rs.min_itm = regression == 0 ? 0u : min_itm;
    rs.max_condition = max_condition;
    rs.early = early;
    rs.expiry = counts[count0 + g.obs - 1];
    rs.ties_training = ties;
    if (first_nosupport != 0xFFFFFFFFu) {
        rs.nosupport_row = static_cast<int32_t>(first_nosupport);
        rs.nosupport_count = nosupport[count0 + first_nosupport];
    } else {
        rs.nosupport_row = -1;
        rs.nosupport_count = 0;
    }
    out[srow] = rs;
}

// Deferred rank/condition gate of the fused method (diagnose_design, regression.rs:540-577): one
// thread per rule accepted during the backward pass. A rule that fails keeps its code for
// k_fit_gate_scan and stops being a regression.
__global__ __launch_bounds__(kBlock) void k_fit_gate(FitParams fp, Model *__restrict__ models, uint32_t count) {
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= count) return;
    if (models[i].kind != kRuleRegression || models[i].gate != kGatePending) return;
    Model m = models[i];
    int64_t index = 0;
    double observed = 0.0, limit = 0.0;
    const int code = gate_fit(m, fp, index, observed, limit);
    if (code != 0) {
        m.kind = kRuleNone;
        m.gate = code;
        m.gate_index = static_cast<int32_t>(index);
        m.gate_observed = observed;
        m.gate_limit = limit;
    } else {
        m.gate = kGateDone;
    }
    models[i] = m;
}

// The earliest failed gate of each strike row in processing order (exercise rows descending)
// becomes the row's regression failure, unless an earlier failure is already recorded.
__global__ __launch_bounds__(kBlock) void k_fit_gate_scan(DevTables t, const uint32_t *__restrict__ rows,
                                                          uint32_t nrows, const Model *__restrict__ models,
                                                          RowError *__restrict__ errors) {
    const uint32_t idx = blockIdx.x * kBlock + threadIdx.x;
    if (idx >= nrows) return;
    const uint32_t srow = rows[idx];
    const DevGroup &g = t.groups[t.strike_group[srow]];
    const uint32_t local = srow - g.strike_begin;
    for (uint32_t step = 0; step < g.fit_rows; ++step) {
        const Model &m = models[g.model_base + local * g.obs + (g.obs - 2 - step)];
        if (m.gate > 0) {
            fail_row(&errors[srow], step, m.gate, m.gate_index, m.gate_observed, m.gate_limit);
            break;
        }
    }
}

// Evaluation normals of every (interval, path) of the evaluation population, time-major
// ([intervals][ne] per group), drawn from each path's stream exactly as k_evaluate draws them; the
// strike tiles of the evaluation pass then read them instead of regenerating Philox and ZIGNOR
// per tile. A path whose draw budget runs out stores NaN from that interval on and reports itself.
__global__ __launch_bounds__(kBlock) void k_eval_normals(DevTables t, const uint2 *__restrict__ items,
                                                         double *__restrict__ znorm,
                                                         unsigned long long *__restrict__ sim_fail) {
    __shared__ double sx[257];
    __shared__ float4 ssq[256];
    __shared__ ExpEntry se[32];
    __shared__ uint64_t ring[kRingSlots * kBlock];
    load_zig_exp(sx, ssq, se);
    const uint2 item = load_item(items, blockIdx.x);
    const DevGroup &g = t.groups[item.x];
    const uint32_t n = g.ne, p = item.y * kEvalChunk + threadIdx.x;
    if (p >= n) return;
    RingStream s;
    s.init(g.eval_seed, g.eval_domain, p, ring, threadIdx.x, kBlock);
    double *out = znorm + g.znorm_base + p;
    bool ok = true;
    for (uint32_t iv = 0; iv < g.intervals; ++iv) {
        double z = 0.0;
        uint32_t draws = 0;
        s.top_up(ok);
        if (ok && !zignor_normal(s, sx, ssq, g_zig_f, FOHIP_ZIG_R, z, draws, se)) ok = false;
        out[static_cast<uint64_t>(iv) * n] = ok ? z : __builtin_nan("");
    }
    if (!ok) atomicMin(sim_fail + item.x, static_cast<unsigned long long>(p));
}

// Serial replay of one failed path (the reference replays the lowest failing index serially to
// report a deterministic error, seeded.rs:155-186): out = {interval, spot, normal, kind}.
__global__ void k_replay_path(DevTables t, uint32_t group, int evaluation, uint32_t path, double *out) {
    __shared__ double sx[257];
    __shared__ float4 ssq[256];
    __shared__ ExpEntry se[32];
    load_zig_exp(sx, ssq, se);
    if (threadIdx.x != 0) return;
    const DevGroup &g = t.groups[group];
    PathStream s;
    s.init(evaluation ? g.eval_seed : g.train_seed, evaluation ? g.eval_domain : g.train_domain, path);
    double spot = g.start;
    out[0] = -1.0;
    for (uint32_t i = 0; i < g.intervals; ++i) {
        double z;
        uint32_t draws = 0;
        if (!zignor_normal(s, sx, ssq, g_zig_f, FOHIP_ZIG_R, z, draws, se)) {
            out[0] = i;
            out[1] = spot;
            out[2] = 0.0;
            out[3] = 1.0;
            return;
        }
        double growth;
        const double next = gbm_advance(spot, t.diffusion[g.step_base + i], t.log_drift[g.step_base + i], z, growth, se);
        if (!isfinite(next)) {
            out[0] = i;
            out[1] = spot;
            out[2] = z;
            out[3] = 0.0;
            return;
        }
        spot = next;
        const uint32_t node = g.node_base + i + 1;
        if (t.node_flags[node] & kNodeCash) spot = apply_cash<double>(g.policy, t.cash[node], spot);
    }
}

// ------------------------------------------------------------------------------ launchers

namespace {
inline uint32_t xcd_blocks(uint32_t nitems, uint32_t columns) {
    return ((nitems + 7u) / 8u) * 8u * columns;
}
}  // namespace

__global__ __launch_bounds__(kBlock) void k_gbm_exp(const double *__restrict__ x, double *__restrict__ out,
                                                    uint32_t n) {
    __shared__ ExpEntry se[32];
    for (int i = threadIdx.x; i < 32; i += kBlock) se[i] = g_exp_table[i];
    __syncthreads();
    const uint32_t i = blockIdx.x * kBlock + threadIdx.x;
    if (i < n) out[i] = x[i] != x[i] ? x[i] : table_exp(x[i], se);   // NaN passes through, as in exp
}

hipError_t launch_gbm_exp(hipStream_t st, const double *x, double *out, uint32_t n) {
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(k_gbm_exp, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, x, out, n);
    return hipGetLastError();
}

hipError_t launch_philox_blocks(hipStream_t st, uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths,
                                uint32_t blocks, uint32_t *out) {
    const uint64_t total = static_cast<uint64_t>(paths) * blocks;
    if (total == 0) return hipSuccess;
    const uint32_t nb = static_cast<uint32_t>((total + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_philox_blocks, dim3(nb), dim3(kBlock), 0, st, seed, domain, path0, paths, blocks, out);
    return hipGetLastError();
}

hipError_t launch_stream_normals(hipStream_t st, uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths,
                                 uint32_t count, double *out, uint32_t *draws, int *fail) {
    if (paths == 0) return hipSuccess;
    hipLaunchKernelGGL(k_stream_normals, dim3((paths + kBlock - 1) / kBlock), dim3(kBlock), 0, st, seed, domain,
                       path0, paths, count, out, draws, fail);
    return hipGetLastError();
}

hipError_t launch_simulate(hipStream_t st, bool f32, const DevTables &t, const uint2 *items, uint32_t nitems,
                           void *paths, int evaluation, unsigned long long *fail) {
    if (nitems == 0) return hipSuccess;
    const dim3 grid(nitems, kChunk / kBlock);
    if (f32)
        hipLaunchKernelGGL(k_simulate<float>, grid, dim3(kBlock), 0, st, t, items, static_cast<float *>(paths),
                           evaluation, fail);
    else
        hipLaunchKernelGGL(k_simulate<double>, grid, dim3(kBlock), 0, st, t, items, static_cast<double *>(paths),
                           evaluation, fail);
    return hipGetLastError();
}

hipError_t launch_fit_prepare(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step,
                              const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                              void *values, const Model *models, double *partials, RowError *errors,
                              unsigned long long *ties) {
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_prepare, dim3(xcd_blocks(nitems, max_strikes)), dim3(kBlock), 0, st, t, fp, step, items,
                       nitems, max_strikes, static_cast<const double *>(paths), static_cast<double *>(values), models,
                       partials, errors, ties);
    return hipGetLastError();
}

hipError_t launch_fit_reduce_welford(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step,
                                     const uint32_t *rows, uint32_t nrows, double *partials, Model *models,
                                     RowError *errors) {
    if (nrows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_reduce_welford, dim3(nrows), dim3(kBlock), 0, st, t, fp, step, rows, partials, models,
                       errors);
    return hipGetLastError();
}

hipError_t launch_fit_moments(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step, int k,
                              const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                              const void *values, const Model *models, double *partials, const RowError *errors) {
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_moments, dim3(xcd_blocks(nitems, max_strikes)), dim3(kBlock), 0, st, t, fp, step, k,
                       items, nitems, max_strikes, static_cast<const double *>(paths),
                       static_cast<const double *>(values), models, partials, errors);
    return hipGetLastError();
}

hipError_t launch_fit_reduce_moments(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step, int k,
                                     const uint32_t *rows, uint32_t nrows, double *partials, Model *models,
                                     RowError *errors) {
    if (nrows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_reduce_moments, dim3(nrows), dim3(kBlock), 0, st, t, fp, step, k, rows, partials, models,
                       errors);
    return hipGetLastError();
}

hipError_t launch_fit_bulk_step(hipStream_t st, bool f32, const DevTables &t, const FitParams &fp, uint32_t step,
                                const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                                void *values, uint8_t *stops, const Model *models, double *partials,
                                RowError *errors) {
    if (nitems == 0) return hipSuccess;
    const uint32_t sets = (max_strikes + kBulkStrikes - 1) / kBulkStrikes;
    const dim3 grid(xcd_blocks(nitems, sets));
    const bool win = fp.tie_window != 0.0;
#define FOHIP_BULK(REAL, WINV)                                                                                  \
    hipLaunchKernelGGL((k_fit_bulk_step<REAL, WINV>), grid, dim3(kBlock), 0, st, t, fp, step, items, nitems, sets, \
                       static_cast<const REAL *>(paths), static_cast<REAL *>(values), models, partials, errors)
#define FOHIP_STOP(REAL, WINV)                                                                                  \
    hipLaunchKernelGGL((k_fit_stop_step<REAL, WINV>), grid, dim3(kBlock), 0, st, t, fp, step, items, nitems, sets, \
                       static_cast<const REAL *>(paths), stops, models, partials, errors)
    if (stops) {
        if (f32) {
            if (win) FOHIP_STOP(float, true);
            else FOHIP_STOP(float, false);
        } else {
            if (win) FOHIP_STOP(double, true);
            else FOHIP_STOP(double, false);
        }
    } else if (f32) {
        if (win) FOHIP_BULK(float, true);
        else FOHIP_BULK(float, false);
    } else {
        if (win) FOHIP_BULK(double, true);
        else FOHIP_BULK(double, false);
    }
#undef FOHIP_BULK
#undef FOHIP_STOP
    return hipGetLastError();
}

hipError_t launch_fit_bulk_reduce(hipStream_t st, const DevTables &t, const FitParams &fp, uint32_t step,
                                  const uint32_t *rows, uint32_t nrows, double *partials, Model *models,
                                  RowError *errors) {
    if (nrows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_bulk_reduce, dim3(nrows), dim3(64), 0, st, t, fp, step, rows, partials, models, errors);
    return hipGetLastError();
}

hipError_t launch_fit_sweep(hipStream_t st, bool f32, const DevTables &t, const FitParams &fp,
                            const SweepPlan &plan, uint32_t slots, const void *paths, void *values,
                            Model *models, RowError *errors) {
    if (slots == 0) return hipSuccess;
    const bool win = fp.tie_window != 0.0;
#define FOHIP_SWEEP(REAL, WINV, QV)                                                                            \
    hipLaunchKernelGGL((k_fit_sweep<REAL, WINV, QV>), dim3(slots), dim3(kSweepBlock), 0, st, t, fp, plan,          \
                       static_cast<const REAL *>(paths), static_cast<REAL *>(values), models, errors)
    if (f32) {
        if (win) FOHIP_SWEEP(float, true, kSweepChunks32);
        else FOHIP_SWEEP(float, false, kSweepChunks32);
    } else {
        if (win) FOHIP_SWEEP(double, true, kSweepChunks64);
        else FOHIP_SWEEP(double, false, kSweepChunks64);
    }
#undef FOHIP_SWEEP
    return hipGetLastError();
}

int sweep_resident_blocks(bool f32, bool window) {
    int n = 0;
    hipError_t e;
    if (f32)
        e = window ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fit_sweep<float, true, kSweepChunks32>, kSweepBlock, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fit_sweep<float, false, kSweepChunks32>, kSweepBlock, 0);
    else
        e = window ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fit_sweep<double, true, kSweepChunks64>, kSweepBlock, 0)
                   : hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, k_fit_sweep<double, false, kSweepChunks64>, kSweepBlock, 0);
    return e == hipSuccess ? n : 0;
}

hipError_t launch_fit_final(hipStream_t st, bool f32, bool fused, const DevTables &t, const FitParams &fp,
                            const uint2 *items, uint32_t nitems, uint32_t max_strikes, const void *paths,
                            const void *values, const uint8_t *stops, const Model *models, double *partials,
                            RowError *errors, unsigned long long *ties) {
    if (nitems == 0) return hipSuccess;
    const dim3 grid(xcd_blocks(nitems, max_strikes));
#define FOHIP_FINAL(REAL, FUSEDV, STOPSV)                                                                          \
    hipLaunchKernelGGL((k_fit_final<REAL, FUSEDV, STOPSV>), grid, dim3(kBlock), 0, st, t, fp, items, nitems,         \
                       max_strikes, static_cast<const REAL *>(paths), static_cast<const REAL *>(values), stops, models, \
                       partials, errors, ties)
    if (f32) {
        if (stops) FOHIP_FINAL(float, true, true);
        else FOHIP_FINAL(float, true, false);
    } else if (fused) {
        if (stops) FOHIP_FINAL(double, true, true);
        else FOHIP_FINAL(double, true, false);
    } else {
        FOHIP_FINAL(double, false, false);
    }
#undef FOHIP_FINAL
    return hipGetLastError();
}

hipError_t launch_fit_reduce_origin(hipStream_t st, const DevTables &t, const FitParams &fp, const uint32_t *rows,
                                    uint32_t nrows, double *partials, Model *models, RowError *errors) {
    if (nrows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_fit_reduce_origin, dim3(nrows), dim3(kBlock), 0, st, t, fp, rows, partials, models, errors);
    return hipGetLastError();
}

// ---------------------------------------------------------------- allocation-free scratch region

// The buffers of an allocation-free pricing live here: kNoAllocScratchBytes per device, loaded with
// the code object and held until the process ends. Host code must never pass `scratch` itself to a
// kernel (in host code the name is a stand-in, not a device address): it uses the address that
// hipGetSymbolAddress returns on each device.
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

hipError_t noalloc_scratch(int device, void **region, size_t *bytes) {
    static std::mutex mu;
    static void *address[kNoAllocMaxDevices] = {};
    if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lock(mu);
    if (!address[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));   // on the current device
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        address[device] = p;
    }
    *region = address[device];
    *bytes = kNoAllocScratchBytes;
    return hipSuccess;
}

uint32_t eval_tile_for(uint32_t max_strikes, bool greeks, uint32_t requested) {
    const uint32_t options_plain[] = {1, 4, 16};
    const uint32_t options_greeks[] = {1, 4, 6, 8};
    const uint32_t *opts = greeks ? options_greeks : options_plain;
    const uint32_t want = requested ? requested : (max_strikes <= 1 ? 1 : (greeks ? 6 : 16));
    uint32_t tile = opts[0];
    const int nopts = greeks ? 4 : 3;
    for (int i = 0; i < nopts; ++i)
        if (opts[i] <= want) tile = opts[i];
    return tile;
}

template <bool FAST, int TS, int G, bool E, bool Z, bool WIN = false>
static hipError_t launch_eval_t(hipStream_t st, const DevTables &t, const FitParams &fp, const uint2 *items,
                                uint32_t nitems, uint32_t max_strikes, uint32_t max_obs, const Model *models,
                                double *partials, uint32_t *counts, uint32_t *nosupport, RowError *errors,
                                unsigned long long *sim_fail, const double *znorm) {
    const uint32_t tiles = (max_strikes + TS - 1) / TS;
    const size_t hist_bytes = static_cast<size_t>(TS) * max_obs * sizeof(uint32_t);
    const int use_lds = hist_bytes <= 32768 ? 1 : 0;
    // staged decision tables for the multi-tile fused forms when they fit beside the histogram;
    // that form keeps 8-bit stopping rows, so it also needs obs <= 255 (implied by the LDS bound)
    const size_t dec_bytes = static_cast<size_t>(TS) * max_obs * kDecDoubles * sizeof(double) + max_obs * sizeof(uint32_t);
    const size_t hist_padded = (hist_bytes + 15) & ~size_t(15);
    if (Z && TS <= 8 && use_lds && max_obs <= 255 && hist_padded + dec_bytes <= 49152) {
        hipLaunchKernelGGL((k_evaluate<FAST, TS, G, E, Z, (Z && TS <= 8), WIN>), dim3(xcd_blocks(nitems, tiles)),
                           dim3(kBlock), hist_padded + dec_bytes, st, t, fp, items, nitems, tiles, models, partials,
                           counts, nosupport, errors, sim_fail, use_lds, znorm);
        return hipGetLastError();
    }
    hipLaunchKernelGGL((k_evaluate<FAST, TS, G, E, Z, false, WIN>), dim3(xcd_blocks(nitems, tiles)), dim3(kBlock),
                       use_lds ? hist_bytes : 0, st, t, fp, items, nitems, tiles, models, partials, counts, nosupport,
                       errors, sim_fail, use_lds, znorm);
    return hipGetLastError();
}

template <bool FAST>
static hipError_t launch_eval_r(hipStream_t st, int gk, bool escrow, uint32_t tile, const DevTables &t, const FitParams &fp,
                                const uint2 *items, uint32_t nitems, uint32_t max_strikes, uint32_t max_obs,
                                const Model *models, double *partials, uint32_t *counts, uint32_t *nosupport,
                                RowError *errors, unsigned long long *sim_fail, const double *znorm) {
    // stored normals serve the fused method's multi-tile launches; a single tile regenerates
#define FOHIP_EVAL_ARGS st, t, fp, items, nitems, max_strikes, max_obs, models, partials, counts, nosupport, errors, \
                        sim_fail, znorm
    // windowed near-tie counts: a diagnostic form, single-strike tiles regenerating their normals
    if (fp.tie_window != 0.0) {
        if (tile != 1 || znorm != nullptr) return hipErrorInvalidValue;
        if (gk)
            return escrow ? launch_eval_t<FAST, 1, 4, true, false, true>(FOHIP_EVAL_ARGS)
                          : launch_eval_t<FAST, 1, 4, false, false, true>(FOHIP_EVAL_ARGS);
        return escrow ? launch_eval_t<FAST, 1, 0, true, false, true>(FOHIP_EVAL_ARGS)
                      : launch_eval_t<FAST, 1, 0, false, false, true>(FOHIP_EVAL_ARGS);
    }
#define FOHIP_EVAL_CASE(TSV, GV, ZV)                                                                               \
    if (tile == TSV && gk == GV && (znorm != nullptr) == ZV)                                                       \
        return escrow ? launch_eval_t<FAST, TSV, GV, true, ZV>(FOHIP_EVAL_ARGS)                                    \
                      : launch_eval_t<FAST, TSV, GV, false, ZV>(FOHIP_EVAL_ARGS);
    // every tile carries only the tangents the requested Greeks need: delta and vega take those of
    // S0 and sigma, the others all four
    FOHIP_EVAL_CASE(1, 2, false)
    FOHIP_EVAL_CASE(4, 2, false)
    FOHIP_EVAL_CASE(6, 2, false)
    FOHIP_EVAL_CASE(8, 2, false)
    if (FAST) {
        FOHIP_EVAL_CASE(4, 2, true)
        FOHIP_EVAL_CASE(6, 2, true)
        FOHIP_EVAL_CASE(8, 2, true)
    }
    FOHIP_EVAL_CASE(1, 0, false)
    FOHIP_EVAL_CASE(4, 0, false)
    FOHIP_EVAL_CASE(16, 0, false)
    FOHIP_EVAL_CASE(1, 4, false)
    FOHIP_EVAL_CASE(4, 4, false)
    FOHIP_EVAL_CASE(6, 4, false)
    FOHIP_EVAL_CASE(8, 4, false)
    if (FAST) {
        FOHIP_EVAL_CASE(4, 0, true)
        FOHIP_EVAL_CASE(16, 0, true)
        FOHIP_EVAL_CASE(4, 4, true)
        FOHIP_EVAL_CASE(6, 4, true)
        FOHIP_EVAL_CASE(8, 4, true)
    }
#undef FOHIP_EVAL_CASE
#undef FOHIP_EVAL_ARGS
    return hipErrorInvalidValue;
}

hipError_t launch_evaluate(hipStream_t st, bool fast, uint32_t greek_mask, bool escrow, uint32_t tile,
                           const DevTables &t, const FitParams &fp, const uint2 *items, uint32_t nitems,
                           uint32_t max_strikes, uint32_t max_obs, const Model *models, double *partials,
                           uint32_t *counts, uint32_t *nosupport, RowError *errors, unsigned long long *sim_fail,
                           const double *znorm) {
    if (nitems == 0) return hipSuccess;
    // tangents carried: none; d/dS0 and d/dsigma (delta and vega, in every dividend model); all four
    const int gk = greek_mask == 0 ? 0 : ((greek_mask & ~(FOHIP_GREEK_DELTA | FOHIP_GREEK_VEGA)) == 0 ? 2 : 4);
    if (fast)
        return launch_eval_r<true>(st, gk, escrow, tile, t, fp, items, nitems, max_strikes, max_obs, models,
                                    partials, counts, nosupport, errors, sim_fail, znorm);
    return launch_eval_r<false>(st, gk, escrow, tile, t, fp, items, nitems, max_strikes, max_obs, models,
                                 partials, counts, nosupport, errors, sim_fail, nullptr);
}

hipError_t launch_fit_gate(hipStream_t st, const DevTables &t, const FitParams &fp, const uint32_t *rows, uint32_t nrows,
                           Model *models, uint32_t model_count, RowError *errors) {
    if (model_count > 0)
        hipLaunchKernelGGL(k_fit_gate, dim3((model_count + kBlock - 1) / kBlock), dim3(kBlock), 0, st, fp, models,
                           model_count);
    if (nrows > 0)
        hipLaunchKernelGGL(k_fit_gate_scan, dim3((nrows + kBlock - 1) / kBlock), dim3(kBlock), 0, st, t, rows, nrows,
                           models, errors);
    return hipGetLastError();
}

hipError_t launch_eval_normals(hipStream_t st, const DevTables &t, const uint2 *items, uint32_t nitems, double *znorm,
                               unsigned long long *sim_fail) {
    if (nitems == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_normals, dim3(nitems), dim3(kBlock), 0, st, t, items, znorm, sim_fail);
    return hipGetLastError();
}

hipError_t launch_row_summary(hipStream_t st, const DevTables &t, const Model *models, const uint32_t *counts,
                              const uint32_t *nosupport, uint32_t rows, RowSummary *out) {
    if (rows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_row_summary, dim3((rows + kWaves - 1) / kWaves), dim3(kBlock), 0, st, t, models, counts,
                       nosupport, rows, out);
    return hipGetLastError();
}

hipError_t launch_replay_path(hipStream_t st, const DevTables &t, uint32_t group, int evaluation, uint32_t path,
                              double *out) {
    hipLaunchKernelGGL(k_replay_path, dim3(1), dim3(64), 0, st, t, group, evaluation, path, out);
    return hipGetLastError();
}

hipError_t launch_eval_reduce(hipStream_t st, const DevTables &t, const uint32_t *rows, uint32_t nrows,
                              const double *partials, int greeks, double *out) {
    if (nrows == 0) return hipSuccess;
    hipLaunchKernelGGL(k_eval_reduce, dim3(nrows), dim3(kBlock), 0, st, t, rows, partials, greeks, out);
    return hipGetLastError();
}

}  // namespace fohip
