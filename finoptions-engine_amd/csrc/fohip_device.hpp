// Device building blocks of the LSM engine for gfx950 (CDNA4, wave64).
//
// Random stream, normal transform and GBM transition restate the reference kernels bit for bit
// where the arithmetic allows it (integer Philox, uniform bit maps, table lookups, FP64 products);
// the exponentials of the GBM step and of the ZIGNOR wedge test are a table-driven binary64 exp
// and the ZIGNOR tail calls the device libm's log; either can differ from glibc by an ulp --
// parity tests count those differences instead of assuming them away.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "exp_table.inc"
#include "fohip_layout.hpp"

namespace fohip {

// ------------------------------------------------------------------ Philox4x32-10 (philox.rs)

constexpr uint32_t kPhiloxM0 = 0xD2511F53u;
constexpr uint32_t kPhiloxM1 = 0xCD9E8D57u;
constexpr uint32_t kPhiloxW0 = 0x9E3779B9u;
constexpr uint32_t kPhiloxW1 = 0xBB67AE85u;

// A wave-uniform word moved into a vector register, where the compiler treats it as varying by
// lane. The Philox keys go through it: the nine bumped keys of a block are loop invariants that
// the compiler hoists out of the path loop, and held in scalar registers (18 of them) they push
// the generation kernels into spilling scalars to vector lanes, reloaded by a vector instruction
// at every use inside the loop; held in vector registers they cost nothing the kernels' LDS-bound
// occupancy does not already leave free.
__device__ __forceinline__ uint32_t in_vgpr(uint32_t x) {
    uint32_t r;
    asm("v_mov_b32 %0, %1" : "=v"(r) : "s"(x));
    return r;
}

// Random123 Philox4x32 with ten rounds (philox.rs:18-41). Round output lanes are
// [hi(M1*c2)^c1^k0, lo(M1*c2), hi(M0*c0)^c3^k1, lo(M0*c0)]; the key is bumped before rounds 2..10.
__device__ __forceinline__ uint4 philox4x32_10(uint32_t c0, uint32_t c1, uint32_t c2, uint32_t c3,
                                               uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int round = 0; round < 10; ++round) {
        if (round) {
            k0 += kPhiloxW0;
            k1 += kPhiloxW1;
        }
        const uint64_t p0 = static_cast<uint64_t>(kPhiloxM0) * c0;
        const uint64_t p1 = static_cast<uint64_t>(kPhiloxM1) * c2;
        // three-input XOR in one v_bitop3_b32 (truth table 0x96), where the compiler would
        // otherwise issue two v_xor_b32
        const uint32_t n0 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p1 >> 32), c1, k0, 0x96);
        const uint32_t n2 = __builtin_amdgcn_bitop3_b32(static_cast<uint32_t>(p0 >> 32), c3, k1, 0x96);
        c1 = static_cast<uint32_t>(p1);
        c3 = static_cast<uint32_t>(p0);
        c0 = n0;
        c2 = n2;
    }
    return make_uint4(c0, c1, c2, c3);
}

// One path-local stream: counter [block_lo, block_hi, path, domain], key [seed_lo, seed_hi]
// (philox.rs:64-113). A u64 draw is word i | word i+1 << 32 and never straddles a block, so each
// block yields exactly two draws; the second is kept for the next call.
struct PathStream {
    uint32_t k0, k1, path, domain;
    uint32_t block_lo, block_hi;
    uint64_t spare;
    bool has_spare;

    __device__ __forceinline__ void init(uint64_t seed, uint32_t domain_id, uint32_t path_id) {
        k0 = in_vgpr(static_cast<uint32_t>(seed));
        k1 = in_vgpr(static_cast<uint32_t>(seed >> 32));
        path = path_id;
        domain = domain_id;
        block_lo = 0;
        block_hi = 0;
        spare = 0;
        has_spare = false;
    }

    __device__ __forceinline__ uint64_t next_u64() {
        if (has_spare) {
            has_spare = false;
            return spare;
        }
        const uint4 w = philox4x32_10(block_lo, block_hi, path, domain, k0, k1);
        block_lo += 1u;
        block_hi += block_lo == 0u ? 1u : 0u;
        spare = static_cast<uint64_t>(w.z) | (static_cast<uint64_t>(w.w) << 32);
        has_spare = true;
        return static_cast<uint64_t>(w.x) | (static_cast<uint64_t>(w.y) << 32);
    }

    __device__ __forceinline__ uint64_t first_u64() { return next_u64(); }
};

// The same stream with its draws staged in a per-lane ring of kRingSlots u64 in LDS. In a wave that
// steps its paths in lockstep, PathStream computes a block whenever any lane's spare is used up;
// after a rejected normal has consumed an extra draw the lanes' spares are out of phase, and the
// wave then computes a block at almost every normal (measured on the stream's own draw counts:
// 99.6 % of normals against the 50 % of lanes in phase). top_up() is called by the whole wave once
// per normal: when some lane holds fewer than kRingLow draws, every lane with room for a block
// receives its next block, one wave-uniform Philox evaluation, which brings the wave down to about
// 0.54 blocks per normal; a lane whose ring runs dry inside a long rejection chain refills on its
// own. The draws are those of PathStream in the same order, bit for bit.
constexpr uint32_t kRingSlots = 8;   // power of two
constexpr uint32_t kRingLow = 3;

struct RingStream {
    uint32_t k0, k1, path, domain;
    uint32_t block_lo, block_hi;
    // draws [head, tail) are in the ring; blocks arrive in pairs of draws, so tail is always even
    // and a block's two draws sit in slots 2k and 2k + 1 (one paired LDS store)
    uint32_t head, tail;
    uint64_t *slot;   // this lane's slots, slot[i * stride]
    uint32_t stride;

    // `lds` holds kRingSlots * stride u64, slot-major, so the lanes of a wave read consecutive
    // words; `lane` indexes this thread's column.
    __device__ __forceinline__ void init(uint64_t seed, uint32_t domain_id, uint32_t path_id, uint64_t *lds,
                                         uint32_t lane, uint32_t columns) {
        k0 = in_vgpr(static_cast<uint32_t>(seed));
        k1 = in_vgpr(static_cast<uint32_t>(seed >> 32));
        path = path_id;
        domain = domain_id;
        block_lo = block_hi = 0;
        head = tail = 0;
        slot = lds + lane;
        stride = columns;
    }

    __device__ __forceinline__ uint32_t count() const { return tail - head; }

    __device__ __forceinline__ void produce() {
        const uint4 w = philox4x32_10(block_lo, block_hi, path, domain, k0, k1);
        block_lo += 1u;
        block_hi += block_lo == 0u ? 1u : 0u;
        uint64_t *d = slot + (tail & (kRingSlots - 2)) * stride;
        d[0] = static_cast<uint64_t>(w.x) | (static_cast<uint64_t>(w.y) << 32);
        d[stride] = static_cast<uint64_t>(w.z) | (static_cast<uint64_t>(w.w) << 32);
        tail += 2;
    }

    // Called by every active lane of the wave at the same point of its loop; `drawing` is false for
    // a lane that draws nothing more (past its population, or failed), which then neither asks
    // for nor receives blocks. Afterwards every drawing lane holds at least two draws.
    __device__ __forceinline__ void top_up(bool drawing) {
        if (__ballot(drawing && count() < kRingLow) != 0ull) {   // wave-uniform
            if (drawing && count() <= kRingSlots - 2) produce();
        }
    }

    __device__ __forceinline__ uint64_t next_u64() {
        if (count() == 0) produce();
        return pop();
    }

    // The first draw after top_up(true): the ring is not empty.
    __device__ __forceinline__ uint64_t first_u64() { return pop(); }

    __device__ __forceinline__ uint64_t pop() {
        const uint64_t v = slot[(head & (kRingSlots - 1)) * stride];
        head += 1;
        return v;
    }
};

// ----------------------------------------------------------------- exponential

// Binary64 exponential of the GBM step and the ZIGNOR wedge test: Tang's table-driven method (P. T. P. Tang, ACM TOMS 15(2),
// 1989) with 32 entries. x = (32 m + j) ln2/32 + r with n = 32 m + j = rint(32 x / ln2) and r by a
// two-part Cody-Waite reduction (|r| <= ln2/64 + tiny); exp(r) - 1 by its Taylor polynomial of
// degree 6 (truncation below 4e-18 relative); 2^(j/32) as a leading + trailing binary64 pair;
// exp(x) = 2^m (T_j + T_j p(r) + t_j) with one final rounding. Within one ulp of the correctly
// rounded value (glibc's exp agrees on 99.5 % of arguments in [-5, 5]). The argument is clamped to
// [-746, 746], where the reduction is exact (|n| < 2^20): the final scaling overflows to +inf above
// ln(DBL_MAX) and rounds into the subnormals below -708.4 (one extra rounding) and to +0 below about
// -745.1, so an argument beyond the clamp gives the same +inf or +0 (and +-inf give +inf and +0).
// The clamp is two min/max instructions where a range test would be a divergent block on every
// call. The argument must not be NaN (the GBM step's and the wedge test's never are: both are
// formed from finite tables and finite normals). No libm call remains, so the loop keeps no libm
// constants in registers.
struct ExpEntry {
    double hi, lo;
};
__device__ __forceinline__ double table_exp(double x, const ExpEntry *__restrict__ tbl) {
    x = fmin(fmax(x, -746.0), 746.0);
    const double n = rint(x * FOHIP_EXP_INV_L);
    double r = fma(-n, FOHIP_EXP_L1, x);
    r = fma(-n, FOHIP_EXP_L2, r);
    const int ni = static_cast<int>(n);
    const ExpEntry e = tbl[ni & 31];
    double q = fma(r, 1.0 / 720.0, 1.0 / 120.0);
    q = fma(r, q, 1.0 / 24.0);
    q = fma(r, q, 1.0 / 6.0);
    q = fma(r, q, 0.5);
    const double p = fma(r * r, q, r);
    return ldexp(e.hi + fma(e.hi, p, e.lo), ni >> 5);
}

// ------------------------------------------------------------ ZIGNOR normal (normal.rs:33-109)

constexpr uint64_t kFractionMask = (uint64_t(1) << 52) - 1;
constexpr uint64_t kExponent0 = uint64_t(1023) << 52;
constexpr uint64_t kExponent1 = uint64_t(1024) << 52;
constexpr int kMaxRejectionIterations = 1048576;
constexpr double kEpsilon = 2.220446049250313e-16;

__device__ __forceinline__ double symmetric_unit(uint64_t bits) {
    return __longlong_as_double(static_cast<long long>(((bits >> 12) & kFractionMask) | kExponent1)) - 3.0;
}
__device__ __forceinline__ double open_unit(uint64_t bits) {
    return __longlong_as_double(static_cast<long long>(((bits >> 12) & kFractionMask) | kExponent0)) -
           (1.0 - kEpsilon / 2.0);
}
__device__ __forceinline__ double standard_unit(uint64_t bits) {
    return static_cast<double>(bits >> 11) * (1.0 / 9007199254740992.0);
}

// Rejected draws: wedge test, redraw, or tail (normal.rs:50-91). Returns false at the bounded
// rejection limit. `draws` counts consumed u64 draws (parity probe only).
// The wedge test "F[i+1] + (F[i] - F[i+1]) u < exp(-value^2 / 2)" is decided first by the layer's
// quadratic squeeze (SQ[i] = {1/w, c, gu, gl}, see gen_zig_tables.py): with t = (|value| - X[i+1]) / w
// and q = t + c t (t - 1), u - q > gu means the test passes and u - q < -gl that it fails, with a
// margin that covers the binary32 arithmetic below (u from the draw's top 23 bits) and every
// rounding of the exact test; only inside the band does the exact test run, reading F from global
// memory. The outcome, and so every normal and draw count, is that of the exact test.
template <typename Stream>
__device__ __forceinline__ bool zignor_rejected(Stream &s, int index, double unit, double value,
                                             const double *__restrict__ X, const float4 *__restrict__ SQ,
                                             const double *__restrict__ F, double zr, double &out,
                                             uint32_t &draws, const ExpEntry *__restrict__ et) {
    for (int it = 0; it < kMaxRejectionIterations; ++it) {
        if (index == 0) {
            for (int tail = 0; tail < kMaxRejectionIterations; ++tail) {
                const double x = log(open_unit(s.next_u64())) / zr;
                const double y = log(open_unit(s.next_u64()));
                draws += 2;
                if (-2.0 * y >= x * x) {
                    out = unit < 0.0 ? x - zr : zr - x;
                    return true;
                }
            }
            return false;
        }
        const uint64_t ubits = s.next_u64();
        draws += 1;
        const float4 sq = SQ[index];
        const float t = static_cast<float>(fabs(value) - X[index + 1]) * sq.x;
        const float q = fmaf(sq.y, t * (t - 1.0f), t);
        // u to 23 bits, truncated: the top 23 bits as the mantissa of a binary32 in [1, 2)
        const float u23 = __uint_as_float(0x3F800000u | (static_cast<uint32_t>(ubits >> 32) >> 9)) - 1.0f;
        const float d = u23 - q;
        bool accept = d > sq.z;
        if (__builtin_expect(!accept && !(d < -sq.w), 0)) {
            const double ordinate = F[index + 1] + (F[index] - F[index + 1]) * standard_unit(ubits);
            accept = ordinate < table_exp(-0.5 * value * value, et);
        }
        if (accept) {
            out = value;
            return true;
        }
        const uint64_t bits = s.next_u64();
        draws += 1;
        index = static_cast<int>(bits & 0xff);
        unit = symmetric_unit(bits);
        value = unit * X[index];
        if (fabs(value) < X[index + 1]) {
            out = value;
            return true;
        }
    }
    return false;
}

template <typename Stream>
__device__ __forceinline__ bool zignor_normal(Stream &s, const double *__restrict__ X,
                                              const float4 *__restrict__ SQ, const double *__restrict__ F,
                                              double zr, double &out, uint32_t &draws,
                                              const ExpEntry *__restrict__ et) {
    const uint64_t bits = s.first_u64();   // a ring stream is topped up before every normal
    draws += 1;
    const int index = static_cast<int>(bits & 0xff);
    const double unit = symmetric_unit(bits);
    const double value = unit * X[index];
    if (fabs(value) < X[index + 1]) {
        out = value;
        return true;
    }
    return zignor_rejected(s, index, unit, value, X, SQ, F, zr, out, draws, et);
}

// ------------------------------------------------------- GBM transition and cash events

// step.rs:89-95: zero is absorbing; S' = S * exp(fma(diffusion, Z, drift)).
__device__ __forceinline__ double gbm_advance(double spot, double diffusion, double drift, double z,
                                              double &growth, const ExpEntry *__restrict__ exp_table) {
    growth = table_exp(fma(diffusion, z, drift), exp_table);
    return spot == 0.0 ? 0.0 : spot * growth;
}
__device__ __forceinline__ float gbm_advance(float spot, float diffusion, float drift, float z,
                                             float &growth) {
    growth = expf(fmaf(diffusion, z, drift));
    return spot == 0.0f ? 0.0f : spot * growth;
}

// dividend.rs:59-81: liquidator max(S - D, 0); survivor S > D ? S - D : S.
template <typename Real>
__device__ __forceinline__ Real apply_cash(int policy, Real amount, Real spot) {
    if (policy == 0) {
        const Real after = spot - amount;
        return after > Real(0) ? after : Real(0);
    }
    return spot > amount ? spot - amount : spot;
}

// ----------------------------------------------------------------- wave-uniform read-only data

// Tables and fitted rules that every lane of a wave addresses identically are read through the
// constant address space, so the compiler emits scalar loads (SGPR results, no per-lane VMEM
// latency chain). Only data that no kernel writes while the reading kernel runs is read this way.
#define FOHIP_UNIFORM __attribute__((address_space(4)))
template <typename T>
__device__ __forceinline__ const FOHIP_UNIFORM T *uniform(const T *p) {
    return (const FOHIP_UNIFORM T *)p;
}

// ------------------------------------------------------------------ regression prediction

// regression.rs:197-224: z = (x - mean) / scale; p_{k+1} = fma(z - alpha_k, p_k, -beta_k p_{k-1});
// prediction = fma(coef_{k+1}, p_{k+1}, prediction).
__device__ __forceinline__ double predict(const Model &m, int degree, double x) {
    const double z = degree == 0 ? 0.0 : (x - m.mean) / m.scale;
    double prediction = m.coef[0], previous = 0.0, current = 1.0;
    for (int k = 0; k < degree; ++k) {
        const double next = fma(z - m.alpha[k], current, -m.beta[k] * previous);
        prediction = fma(m.coef[k + 1], next, prediction);
        previous = current;
        current = next;
    }
    return prediction;
}
__device__ __forceinline__ float predict(const Model &m, int degree, float x) {
    const float z = degree == 0 ? 0.0f : (x - static_cast<float>(m.mean)) / static_cast<float>(m.scale);
    float prediction = static_cast<float>(m.coef[0]), previous = 0.0f, current = 1.0f;
    for (int k = 0; k < degree; ++k) {
        const float next =
            fmaf(z - static_cast<float>(m.alpha[k]), current, -static_cast<float>(m.beta[k]) * previous);
        prediction = fmaf(static_cast<float>(m.coef[k + 1]), next, prediction);
        previous = current;
        current = next;
    }
    return prediction;
}

// The same recurrence read through any pointer kind (generic or uniform).
template <typename MP>
__device__ __forceinline__ double predict_ref(MP m, int degree, double x) {
    const double z = degree == 0 ? 0.0 : (x - m->mean) / m->scale;
    double prediction = m->coef[0], previous = 0.0, current = 1.0;
    for (int k = 0; k < degree; ++k) {
        const double next = fma(z - m->alpha[k], current, -m->beta[k] * previous);
        prediction = fma(m->coef[k + 1], next, prediction);
        previous = current;
        current = next;
    }
    return prediction;
}

// Evaluation form of an accepted rule at spot S: z = fma(S, 1/(scale K), -mean/scale), Horner in
// z. Unused degrees carry zero coefficients, so the fixed four FMAs serve every degree.
// c * z + u with u wave-uniform: the VOP3 form reads u from a scalar register directly (the
// accumulating VOP2 form the compiler otherwise picks first copies u into the result register).
__device__ __forceinline__ double fma_uniform_addend(double c, double z, double u) {
    double r;
    asm("v_fma_f64 %0, %1, %2, %3" : "=v"(r) : "v"(c), "v"(z), "s"(u));
    return r;
}

template <typename MP>
__device__ __forceinline__ double continuation_fast(MP m, double spot) {
    const double z = fma(spot, m->a, m->b);
    double c = fma(m->hz[4], z, m->hz[3]);
    c = fma_uniform_addend(c, z, m->hz[2]);
    c = fma_uniform_addend(c, z, m->hz[1]);
    return fma_uniform_addend(c, z, m->hz[0]);
}

// ExerciseTieRule::should_exercise (policy.rs:21-26): strictly greater, or greater or equal. Written
// as two comparisons joined by lane-mask logic rather than a select between them.
__device__ __forceinline__ bool should_exercise(int tie, double intrinsic, double continuation) {
    return (intrinsic > continuation) | ((tie != 0) & (intrinsic == continuation));
}

// Tangents of the observed stock S = F + A(node) w.r.t. (S0, sigma, r, q) from those of the
// simulated state (d/dF0, d/dsigma_F, drift d/dr, d/dq), by the chain rule through F0 = S0 - A0(r)
// and sigma_F(S0, sigma, r), plus dA(node)/dr. The spot models have c = (0, 1, 0, 0) and dA = 0,
// which returns the state tangents unchanged.
__device__ __forceinline__ void observed_tangents(const double (&tan)[4], const FOHIP_UNIFORM double *c, double dA,
                                                  double (&o)[4]) {
    o[0] = tan[0] + c[0] * tan[1];
    o[1] = c[1] * tan[1];
    o[2] = tan[2] + c[2] * tan[0] + c[3] * tan[1] + dA;
    o[3] = tan[3];
}

// ------------------------------------------------------- deterministic block reductions

constexpr int kBlock = 256;      // threads per block
constexpr int kWaves = kBlock / 64;

// Wave total through DPP lane exchanges: pairs, quads, half-rows and rows of 16 are combined
// with commutative adds (both partners hold identical sums), then the four row totals are
// combined as (r0 + r1) + (r2 + r3). The order is fixed by lane indices, so the result is
// reproducible; every lane receives the total.
template <int CTRL>
__device__ __forceinline__ double dpp_partner(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_total(double v) {
    v += dpp_partner<0xB1>(v);    // quad_perm [1,0,3,2]
    v += dpp_partner<0x4E>(v);    // quad_perm [2,3,0,1]
    v += dpp_partner<0x141>(v);   // row_half_mirror
    v += dpp_partner<0x140>(v);   // row_mirror
    return (readlane_d(v, 0) + readlane_d(v, 16)) + (readlane_d(v, 32) + readlane_d(v, 48));
}

// Block sum of N values per thread; the result is returned in thread 0 (fixed order
// ((w0 + w1) + (w2 + w3)) over the wave totals). `scratch` holds kWaves * N doubles. Every thread
// of the block must call it.
template <int N>
__device__ __forceinline__ void block_sum(double (&v)[N], double *scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
    for (int i = 0; i < N; ++i) v[i] = wave_total(v[i]);
    if (lane == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i) scratch[wave * N + i] = v[i];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int i = 0; i < N; ++i)
            v[i] = (scratch[0 * N + i] + scratch[1 * N + i]) + (scratch[2 * N + i] + scratch[3 * N + i]);
    }
    __syncthreads();
}

// Wave sums of N <= 16 doubles per lane through a transpose in LDS: every lane stores its N values,
// then lane 4q + j (j < 4) adds entries 16j .. 16j + 15 of value q in index order and the quad
// combines its four partials by two commutative exchanges. Lane 4q returns the total of value q
// (other lanes return unspecified values). `buf` is this wave's private area of N * 66 doubles.
// The order of every addition is fixed by lane indices, so the sums are reproducible.
template <int N>
__device__ __forceinline__ double wave_sums_lds(const double (&v)[N], double *buf) {
    static_assert(N <= 16, "at most 16 values");
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int q = 0; q < N; ++q) buf[q * 66 + lane] = v[q];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    const int q = lane >> 2, part = lane & 3;
    double sum = 0.0;
    if (q < N) {
        const double2 *row = reinterpret_cast<const double2 *>(buf + q * 66 + part * 16);
#pragma unroll
        for (int k = 0; k < 8; ++k) {
            const double2 pair = row[k];
            sum += pair.x;
            sum += pair.y;
        }
    }
    sum += dpp_partner<0xB1>(sum);   // quad_perm [1,0,3,2]
    sum += dpp_partner<0x4E>(sum);   // quad_perm [2,3,0,1]
    __builtin_amdgcn_wave_barrier();   // buf may be rewritten by the next call
    return sum;
}

// Chan et al. pairwise merge of (count, mean, M2) with a formula symmetric in its operands.
struct Moments {
    double n, mean, m2;
};
__device__ __forceinline__ Moments chan_merge(Moments a, Moments b) {
    const double n = a.n + b.n;
    if (n == 0.0) return {0.0, 0.0, 0.0};
    if (a.n == 0.0) return b;
    if (b.n == 0.0) return a;
    const double delta = b.mean - a.mean;
    Moments r;
    r.n = n;
    r.mean = (a.n * a.mean + b.n * b.mean) / n;
    r.m2 = (a.m2 + b.m2) + delta * delta * (a.n * b.n / n);
    return r;
}

__device__ __forceinline__ Moments wave_chan(Moments m) {
#pragma unroll
    for (int offset = 32; offset > 0; offset >>= 1) {
        Moments o;
        o.n = __shfl_down(m.n, offset, 64);
        o.mean = __shfl_down(m.mean, offset, 64);
        o.m2 = __shfl_down(m.m2, offset, 64);
        m = chan_merge(m, o);
    }
    return m;
}

// Block Chan merge; result valid in thread 0. `scratch` holds 3 * kWaves doubles.
__device__ __forceinline__ Moments block_chan(Moments m, double *scratch) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    m = wave_chan(m);
    if (lane == 0) {
        scratch[wave * 3 + 0] = m.n;
        scratch[wave * 3 + 1] = m.mean;
        scratch[wave * 3 + 2] = m.m2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        Moments w[kWaves];
        for (int i = 0; i < kWaves; ++i) w[i] = {scratch[i * 3], scratch[i * 3 + 1], scratch[i * 3 + 2]};
        m = chan_merge(chan_merge(w[0], w[1]), chan_merge(w[2], w[3]));
    }
    __syncthreads();
    return m;
}

// Per-thread Welford update of regression.rs:288-293.
__device__ __forceinline__ void welford_push(Moments &m, double x) {
    m.n += 1.0;
    const double delta = x - m.mean;
    m.mean += delta / m.n;
    m.m2 = fma(delta, x - m.mean, m.m2);
}

}  // namespace fohip
