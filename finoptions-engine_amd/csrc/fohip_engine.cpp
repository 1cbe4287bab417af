// Host side of libfohip: request validation, event-grid and curve construction, HBM layout,
// launch orchestration and the C ABI declared in include/fohip.h.
//
// The pricing pipeline per batch of groups is:
//   training simulation -> for each backward step: prepare, support/scale reduce,
//   (d + 1) x [moments, Stieltjes reduce] (FP64) or [Gram, LDL^T reduce] (FP32 bulk)
//   -> final discount + origin decision -> fused evaluation -> per-row summary.
// Everything is enqueued on the executor's stream; the host synchronizes once per batch.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "fohip.h"
#include "fohip_kernels.hpp"
#include "fohip_layout.hpp"

using namespace fohip;

// The fused fit step's unmasked-load test at the largest path counts the engine accepts: the last
// (ragged) chunk of 2^32 - 1 paths is chunk 2^21 - 1, whose end 2^32 lies past the population.
static_assert(bulk_chunk_is_full(0, kBulkChunk) && !bulk_chunk_is_full(0, kBulkChunk - 1), "chunk bound");
static_assert(!bulk_chunk_is_full((1u << 21) - 1, 0xFFFFFFFFu), "chunk bound: 32-bit wrap at 2^32 - 1 paths");
static_assert(bulk_chunk_is_full((1u << 21) - 2, 0xFFFFFFFFu), "chunk bound: next-to-last chunk is full");
static_assert(bulk_chunk_is_full((1u << 20) - 1, 1u << 31), "chunk bound: exact multiple");

// ================================================================================ errors

namespace {

int fail(fohip_error *err, int code, int phase, const char *fmt, ...) {
    if (err) {
        err->code = code;
        err->phase = phase;
        va_list ap;
        va_start(ap, fmt);
        vsnprintf(err->message, sizeof err->message, fmt, ap);
        va_end(ap);
    }
    return code;
}

void clear_error(fohip_error *err) {
    if (!err) return;
    std::memset(err, 0, sizeof *err);
    err->row = err->node = err->path = err->index = -1;
}

void set_fields(fohip_error *err, int64_t row, int64_t node, int64_t path, int64_t index,
                double observed, double limit) {
    if (!err) return;
    err->row = row;
    err->node = node;
    err->path = path;
    err->index = index;
    err->observed = observed;
    err->limit = limit;
}

bool canonical_zero_or_positive(double v) {
    return std::isfinite(v) && v >= 0.0 && !(v == 0.0 && std::signbit(v));
}

// A binary64 value as the reference's messages print it (Rust's `{:?}`): the shortest digits that
// round-trip, a decimal point on integral values ("1.0"), scientific form outside [1e-4, 1e16).
struct F64Text {
    char s[40];
};
F64Text rust_f64(double v) {
    F64Text out{};
    if (!std::isfinite(v)) {
        std::snprintf(out.s, sizeof out.s, "%s", std::isnan(v) ? "NaN" : (v > 0 ? "inf" : "-inf"));
        return out;
    }
    if (v == 0.0) {
        std::snprintf(out.s, sizeof out.s, "%s", std::signbit(v) ? "-0.0" : "0.0");
        return out;
    }
    char digits[40];
    int prec = 0;
    for (; prec < 17; ++prec) {   // shortest round-trip mantissa
        std::snprintf(digits, sizeof digits, "%.*e", prec, v);
        if (std::strtod(digits, nullptr) == v) break;
    }
    const char *e = std::strchr(digits, 'e');
    const int exp10 = std::atoi(e + 1);
    std::string mant(static_cast<const char *>(digits), e);   // "-d.ddd"
    const bool neg = mant[0] == '-';
    std::string ds;
    for (char c : mant)
        if (c >= '0' && c <= '9') ds.push_back(c);
    while (ds.size() > 1 && ds.back() == '0') ds.pop_back();
    std::string body;
    if (exp10 < -4 || exp10 >= 16) {
        body = ds.substr(0, 1);
        if (ds.size() > 1) body += "." + ds.substr(1);
        body += "e" + std::to_string(exp10);
    } else if (exp10 < 0) {
        body = "0." + std::string(static_cast<size_t>(-exp10 - 1), '0') + ds;
    } else {
        const size_t point = static_cast<size_t>(exp10) + 1;
        if (ds.size() <= point) body = ds + std::string(point - ds.size(), '0') + ".0";
        else body = ds.substr(0, point) + "." + ds.substr(point);
    }
    std::snprintf(out.s, sizeof out.s, "%s%s", neg ? "-" : "", body.c_str());
    return out;
}

#define HIP_TRY(expr)                                                                          \
    do {                                                                                       \
        hipError_t _e = (expr);                                                                \
        if (_e != hipSuccess)                                                                  \
            return fail(err, _e == hipErrorOutOfMemory ? FOHIP_E_OUT_OF_MEMORY : FOHIP_E_HIP,  \
                        FOHIP_PHASE_NONE, "HIP call %s failed: %s", #expr,                     \
                        hipGetErrorString(_e));                                                \
    } while (0)

}  // namespace

// ============================================================================ validation

extern "C" int fohip_validate_time_grid(const double *times, size_t nodes, fohip_error *err) {
    clear_error(err);
    // time.rs:31-64
    if (nodes == 0 || !times) return fail(err, FOHIP_E_INVALID_TIME_GRID, 0, "time grid must contain at least the origin");
    if (!std::isfinite(times[0]))
        return fail(err, FOHIP_E_INVALID_TIME_GRID, 0, "time at node 0 must be finite; observed %.17g", times[0]);
    if (times[0] != 0.0 || std::signbit(times[0]))
        return fail(err, FOHIP_E_INVALID_TIME_GRID, 0, "time grid must start at exact canonical 0.0; observed %.17g",
                    times[0]);
    if (nodes - 1 > 0xFFFFFFFFull) return fail(err, FOHIP_E_INVALID_TIME_GRID, 0, "time grid has too many nodes");
    for (size_t i = 1; i < nodes; ++i) {
        if (!std::isfinite(times[i])) {
            set_fields(err, -1, (int64_t)i, -1, -1, times[i], 0);
            return fail(err, FOHIP_E_INVALID_TIME_GRID, 0, "time at node %zu must be finite; observed %.17g", i, times[i]);
        }
        if (times[i] <= times[i - 1]) {
            set_fields(err, -1, (int64_t)i, -1, -1, times[i], times[i - 1]);
            return fail(err, FOHIP_E_INVALID_TIME_GRID, 0,
                        "time grid must be strictly increasing; node %zu = %.17g follows node %zu = %.17g", i,
                        times[i], i - 1, times[i - 1]);
        }
    }
    return FOHIP_OK;
}

extern "C" int fohip_validate_log_discount_curve(const double *times, const double *logd, size_t nodes,
                                                 fohip_error *err) {
    int rc = fohip_validate_time_grid(times, nodes, err);
    if (rc) return rc;
    // curve.rs:24-62
    if (logd[0] != 0.0 || std::signbit(logd[0]))
        return fail(err, FOHIP_E_INVALID_LOG_DISCOUNT_CURVE, 0,
                    "log discount at the origin must be exact canonical 0.0; observed %.17g", logd[0]);
    for (size_t i = 0; i < nodes; ++i) {
        if (!std::isfinite(logd[i])) {
            set_fields(err, -1, (int64_t)i, -1, -1, logd[i], 0);
            return fail(err, FOHIP_E_INVALID_LOG_DISCOUNT_CURVE, 0, "log discount at node %zu must be finite; observed %.17g",
                        i, logd[i]);
        }
        if (i == 0) continue;
        const double interval = logd[i] - logd[i - 1];
        const double d = std::exp(interval);
        if (!std::isfinite(d) || d <= 0.0)
            return fail(err, FOHIP_E_INVALID_LOG_DISCOUNT_CURVE, 0,
                        "discount from node %zu to %zu is not representable; interval log discount %.17g", i - 1, i,
                        interval);
    }
    return FOHIP_OK;
}

extern "C" int fohip_validate_cumulative_yield_curve(const double *times, const double *q, size_t nodes,
                                                     fohip_error *err) {
    int rc = fohip_validate_time_grid(times, nodes, err);
    if (rc) return rc;
    // yield_curve.rs:28-80
    if (q[0] != 0.0 || std::signbit(q[0]))
        return fail(err, FOHIP_E_INVALID_YIELD_CURVE, 0,
                    "cumulative dividend yield at the origin must be exact canonical 0.0; observed %.17g", q[0]);
    for (size_t i = 0; i < nodes; ++i) {
        if (!canonical_zero_or_positive(q[i]))
            return fail(err, FOHIP_E_INVALID_YIELD_CURVE, 0,
                        "cumulative dividend yield at node %zu must be finite and nonnegative; observed %.17g", i, q[i]);
        if (i == 0) continue;
        if (q[i] < q[i - 1])
            return fail(err, FOHIP_E_INVALID_YIELD_CURVE, 0,
                        "cumulative dividend yield decreases from node %zu (%.17g) to node %zu (%.17g)", i - 1, q[i - 1],
                        i, q[i]);
        const double payout = std::exp(-(q[i] - q[i - 1]));
        if (!std::isfinite(payout) || payout <= 0.0)
            return fail(err, FOHIP_E_INVALID_YIELD_CURVE, 0,
                        "dividend-yield payout discount from node %zu to %zu is not representable", i - 1, i);
    }
    return FOHIP_OK;
}

extern "C" int fohip_validate_cash_dividend_amount(double amount, fohip_error *err) {
    clear_error(err);
    if (!canonical_zero_or_positive(amount))
        return fail(err, FOHIP_E_INVALID_CASH_DIVIDEND, 0,
                    "cash dividend amount must be finite, nonnegative and canonical; observed %.17g", amount);
    return FOHIP_OK;
}

extern "C" int fohip_validate_constant_volatility(double v, fohip_error *err) {
    clear_error(err);
    if (!std::isfinite(v) || v <= 0.0)
        return fail(err, FOHIP_E_INVALID_VOLATILITY, 0,
                    "constant volatility must be finite and strictly positive; observed %.17g", v);
    return FOHIP_OK;
}

extern "C" int fohip_validate_event_schedule(const double *times, size_t nodes, const uint32_t *dnodes,
                                             const double *damounts, size_t dividends, const uint32_t *enodes,
                                             size_t exercise_count, uint32_t expiry, fohip_error *err) {
    int rc = fohip_validate_time_grid(times, nodes, err);
    if (rc) return rc;
    // event.rs:42-112
    const uint64_t expected = nodes - 1;
    if (expected == 0 || expiry != expected)
        return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "expiry must be the final node %llu; observed %u",
                    (unsigned long long)expected, expiry);
    for (size_t i = 0; i < dividends; ++i) {
        if (dnodes[i] == 0 || dnodes[i] > expiry)
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "cash dividend node %u must lie in 1..=%u", dnodes[i], expiry);
        if (i > 0 && dnodes[i - 1] == dnodes[i])
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "duplicate cash dividend node %u", dnodes[i]);
        if (i > 0 && dnodes[i - 1] > dnodes[i])
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "cash dividend nodes are unordered: %u then %u", dnodes[i - 1],
                        dnodes[i]);
        rc = fohip_validate_cash_dividend_amount(damounts[i], err);
        if (rc) return rc;
    }
    for (size_t i = 0; i < exercise_count; ++i) {
        if (enodes[i] >= expiry)
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "exercise node %u must precede expiry %u", enodes[i], expiry);
        if (i > 0 && enodes[i - 1] == enodes[i])
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "duplicate exercise node %u", enodes[i]);
        if (i > 0 && enodes[i - 1] > enodes[i])
            return fail(err, FOHIP_E_INVALID_SCHEDULE, 0, "exercise nodes are unordered: %u then %u", enodes[i - 1],
                        enodes[i]);
    }
    return FOHIP_OK;
}

extern "C" int fohip_validate_ragged_offsets(const int64_t *offsets, size_t len, size_t values_len, fohip_error *err) {
    clear_error(err);
    // batch/src/view.rs:99-160: at least two offsets, nonnegative and nondecreasing; offsets may
    // start above zero (a sliced Arrow list) and then address child[offset - offsets[0]].
    if (len < 2 || !offsets) return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "ragged offsets need at least two entries");
    for (size_t i = 0; i < len; ++i) {
        if (offsets[i] < 0)
            return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "ragged offset %zu is negative: %lld", i, (long long)offsets[i]);
        if (i > 0 && offsets[i] < offsets[i - 1])
            return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "ragged offsets decrease at %zu", i);
    }
    if (static_cast<uint64_t>(offsets[len - 1] - offsets[0]) != values_len)
        return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "ragged offsets span %lld child values but %zu are present",
                    (long long)(offsets[len - 1] - offsets[0]), values_len);
    return FOHIP_OK;
}

extern "C" int fohip_validate_ragged_exercise_times(const int64_t *offsets, size_t len, const double *times,
                                                    fohip_error *err) {
    clear_error(err);
    if (len < 2 || !offsets) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "ragged offsets need at least two entries");
    const int64_t base = offsets[0];
    for (size_t r = 0; r + 1 < len; ++r) {
        const double *t = times + (offsets[r] - base);
        const size_t n = static_cast<size_t>(offsets[r + 1] - offsets[r]);
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(t[i]) || t[i] <= 0.0)
                return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "row %zu exercise times must be finite and positive", r);
        for (size_t i = 1; i < n; ++i)
            if (!(t[i] > t[i - 1]))
                return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "row %zu exercise times must be strictly increasing", r);
    }
    return FOHIP_OK;
}

extern "C" int fohip_validate_ragged_cash_dividends(const int64_t *offsets, size_t len, const double *times,
                                                    const double *amounts, fohip_error *err) {
    clear_error(err);
    if (len < 2 || !offsets) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "ragged offsets need at least two entries");
    const int64_t base = offsets[0];
    for (size_t r = 0; r + 1 < len; ++r) {
        const double *t = times + (offsets[r] - base), *a = amounts + (offsets[r] - base);
        const size_t n = static_cast<size_t>(offsets[r + 1] - offsets[r]);
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(t[i]) || t[i] <= 0.0)
                return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "row %zu cash-dividend times must be finite and positive",
                            r);
        for (size_t i = 1; i < n; ++i)
            if (t[i] < t[i - 1])
                return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "row %zu cash-dividend times must be nondecreasing", r);
        for (size_t i = 0; i < n; ++i)
            if (!std::isfinite(a[i]) || a[i] < 0.0 || (a[i] == 0.0 && std::signbit(a[i])))
                return fail(err, FOHIP_E_INVALID_ARGUMENT, 0,
                            "row %zu cash-dividend amounts must be finite and nonnegative", r);
    }
    return FOHIP_OK;
}

// ============================================================================== executor

namespace {

struct DevBuf {
    void *ptr = nullptr;
    size_t cap = 0;
    // Grows the buffer. A batch queued behind another may still be reading the old buffer, so the
    // device drains before it is released (growth happens on the first requests of a shape only).
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (ptr) {
            hipError_t e = hipDeviceSynchronize();
            if (e == hipSuccess) e = hipFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, 256);
        hipError_t e = hipMalloc(&ptr, want);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    template <typename T> T *as() const { return static_cast<T *>(ptr); }
};

// Page-locked host memory: copies from and to it are DMA transfers that return at once, where a
// pageable buffer costs a staged, blocking copy per call. Grown only between batches, after the
// stream has drained.
struct PinnedBuf {
    char *ptr = nullptr;
    size_t cap = 0;
    hipError_t ensure(size_t bytes) {
        if (bytes <= cap) return hipSuccess;
        if (ptr) {
            hipError_t e = hipHostFree(ptr);
            if (e != hipSuccess) return e;
            ptr = nullptr;
            cap = 0;
        }
        size_t want = std::max<size_t>(bytes, size_t(1) << 16);
        hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&ptr), want, hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = want;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

// The device and host buffers of one batch in flight, and the events that order its phases across
// the executor's two streams: simulation done (generation stream), backward induction done
// (fit stream), results downloaded (generation stream).
struct BatchSlot {
    DevBuf paths, values, stops, models, part_t, part_e, counts, nosup, errors, fail_t, fail_e, tables, eval_out,
        summary, replay, znorm, ties, sweep_sync, sweep_part;
    PinnedBuf staging;   // the batch's tables and items on the way in, results on the way out
    hipEvent_t simulated = nullptr, fitted = nullptr, done = nullptr;
    void release() {
        for (DevBuf *b : {&paths, &values, &stops, &models, &part_t, &part_e, &counts, &nosup, &errors, &fail_t,
                          &fail_e, &tables, &eval_out, &summary, &replay, &znorm, &ties, &sweep_sync, &sweep_part})
            b->release();
        staging.release();
        for (hipEvent_t e : {simulated, fitted, done})
            if (e) (void)hipEventDestroy(e);
        simulated = fitted = done = nullptr;
    }
};

}  // namespace

struct fohip_executor {
    int32_t device = 0;
    int32_t precision = 0;
    hipStream_t stream = nullptr;   // backward induction; every single-stream entry point
    hipStream_t gen = nullptr;      // uploads, simulation, evaluation and downloads of batches
    char name[256] = {0};
    fohip_tuning tuning = {0, 0, 0, 0, 0};
    double near_tie_window = 0.0;   // 0: exact ties; w > 0: |intrinsic - continuation| <= w |continuation|
    size_t batch_budget = 0;   // default bytes for paths + values of a batch (serial schedule)
    std::mutex mu;
    BatchSlot slot[2];
    DevBuf items_t, probe_a, probe_b, probe_c;
    // blocks of the fit sweep the device holds at once, per (binary32 storage, windowed ties);
    // -1 until first asked
    int64_t sweep_slots[2][2] = {{-1, -1}, {-1, -1}};
    // host copies from the last grouped call, kept for fohip_grouped_diagnostics
    std::vector<Model> last_models;
    std::vector<uint32_t> last_model_base, last_obs, last_strike_begin, last_strike_end;
    std::vector<uint32_t> last_row_nodes;   // per group: observation nodes concatenated
    std::vector<uint32_t> last_row_base;
    // kernel timing
    bool profiling = false;
    struct Pending {
        const char *name;
        hipEvent_t start, stop;
    };
    std::vector<Pending> pending;
    std::vector<hipEvent_t> event_pool;
    std::vector<fohip_kernel_stat> stats;
    hipEvent_t take_event() {
        if (!event_pool.empty()) {
            hipEvent_t e = event_pool.back();
            event_pool.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        return e;
    }
    // Accumulates the launches whose stop event has completed; the others stay pending (a batch in
    // flight on the other stream).
    void collect() {
        std::vector<Pending> later;
        for (auto &p : pending) {
            if (hipEventQuery(p.stop) != hipSuccess) {
                later.push_back(p);
                continue;
            }
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, p.start, p.stop) == hipSuccess) {
                fohip_kernel_stat *st = nullptr;
                for (auto &s : stats)
                    if (std::strncmp(s.name, p.name, sizeof s.name) == 0) st = &s;
                if (!st) {
                    fohip_kernel_stat ns;
                    std::memset(&ns, 0, sizeof ns);
                    std::snprintf(ns.name, sizeof ns.name, "%s", p.name);
                    stats.push_back(ns);
                    st = &stats.back();
                }
                st->launches += 1;
                st->total_ms += ms;
            }
            event_pool.push_back(p.start);
            event_pool.push_back(p.stop);
        }
        pending.swap(later);
    }
};

namespace {
// Brackets one launch with events on the stream it is queued on when profiling is enabled.
struct ProfScope {
    fohip_executor *ex;
    hipStream_t stream;
    const char *name;
    hipEvent_t start = nullptr;
    ProfScope(fohip_executor *e, hipStream_t s, const char *n) : ex(e), stream(s), name(n) {
        if (ex->profiling) {
            start = ex->take_event();
            if (start) (void)hipEventRecord(start, stream);
        }
    }
    ~ProfScope() {
        if (ex->profiling && start) {
            hipEvent_t stop = ex->take_event();
            if (stop) {
                (void)hipEventRecord(stop, stream);
                ex->pending.push_back({name, start, stop});
            }
        }
    }
};
}  // namespace

#define PROF_LAUNCH(stream, name, expr)          \
    do {                                         \
        ProfScope _scope(ex, stream, name);      \
        HIP_TRY(expr);                           \
    } while (0)

extern "C" int fohip_abi_version(void) { return FOHIP_ABI_VERSION; }

extern "C" int fohip_device_count(int32_t *count) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess) n = 0;
    if (count) *count = n;
    return FOHIP_OK;
}

extern "C" int fohip_executor_create(int32_t device, int32_t precision, fohip_executor **out, fohip_error *err) {
    clear_error(err);
    if (!out) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "executor output pointer is null");
    *out = nullptr;
    if (precision != FOHIP_PRECISION_F64 && precision != FOHIP_PRECISION_F32_BULK &&
        precision != FOHIP_PRECISION_FUSED_F64)
        return fail(err, FOHIP_E_INVALID_ARGUMENT, 0,
                    "precision must be FP64 (0), FP32 bulk (1) or fused FP64 (2); observed %d", precision);
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= 0)
        return fail(err, FOHIP_E_NO_DEVICE, 0, "no HIP device is available; the HIP executor has no fallback");
    if (device < 0 || device >= n)
        return fail(err, FOHIP_E_NO_DEVICE, 0, "HIP device %d does not exist (%d visible)", device, n);
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    auto *ex = new (std::nothrow) fohip_executor();
    if (!ex) return fail(err, FOHIP_E_OUT_OF_MEMORY, 0, "executor allocation failed");
    ex->device = device;
    ex->precision = precision;
    std::snprintf(ex->name, sizeof ex->name, "%s (%s)", prop.name, prop.gcnArchName);
    // Paths + values of one batch of groups stay below a quarter of device memory (the two slots
    // of the overlapped schedule hold an eighth each).
    ex->batch_budget = std::max<size_t>(prop.totalGlobalMem / 4, size_t(1) << 30);
    // both streams at the default priority
    e = hipStreamCreateWithFlags(&ex->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipStreamCreateWithFlags(&ex->gen, hipStreamNonBlocking);
    for (BatchSlot &b : ex->slot)
        for (hipEvent_t *ev : {&b.simulated, &b.fitted, &b.done})
            if (e == hipSuccess) e = hipEventCreateWithFlags(ev, hipEventDisableTiming);
    if (e != hipSuccess) {
        (void)fohip_executor_destroy(ex);
        return fail(err, FOHIP_E_HIP, 0, "stream or event creation failed: %s", hipGetErrorString(e));
    }
    *out = ex;
    return FOHIP_OK;
}

extern "C" int fohip_executor_destroy(fohip_executor *ex) {
    if (!ex) return FOHIP_OK;
    (void)hipSetDevice(ex->device);
    for (hipStream_t s : {ex->stream, ex->gen})
        if (s) (void)hipStreamSynchronize(s);
    for (BatchSlot &b : ex->slot) b.release();
    for (DevBuf *b : {&ex->items_t, &ex->probe_a, &ex->probe_b, &ex->probe_c}) b->release();
    for (auto &p : ex->pending) {
        (void)hipEventDestroy(p.start);
        (void)hipEventDestroy(p.stop);
    }
    for (hipEvent_t e : ex->event_pool) (void)hipEventDestroy(e);
    for (hipStream_t s : {ex->stream, ex->gen})
        if (s) (void)hipStreamDestroy(s);
    delete ex;
    return FOHIP_OK;
}

extern "C" int fohip_executor_info(const fohip_executor *ex, int32_t *device, int32_t *precision, char *name,
                                   size_t cap) {
    if (!ex) return FOHIP_E_INVALID_ARGUMENT;
    if (device) *device = ex->device;
    if (precision) *precision = ex->precision;
    if (name && cap) std::snprintf(name, cap, "%s", ex->name);
    return FOHIP_OK;
}

extern "C" int fohip_set_profiling(fohip_executor *ex, int32_t enabled) {
    if (!ex) return FOHIP_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ex->mu);
    ex->profiling = enabled != 0;
    return FOHIP_OK;
}

extern "C" int fohip_get_profile(fohip_executor *ex, fohip_kernel_stat *out, size_t cap, size_t *count) {
    if (!ex) return FOHIP_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ex->mu);
    if (count) *count = ex->stats.size();
    for (size_t i = 0; i < ex->stats.size() && i < cap && out; ++i) out[i] = ex->stats[i];
    return FOHIP_OK;
}

extern "C" int fohip_reset_profile(fohip_executor *ex) {
    if (!ex) return FOHIP_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ex->mu);
    ex->stats.clear();
    return FOHIP_OK;
}

extern "C" int fohip_set_tuning(fohip_executor *ex, const fohip_tuning *t) {
    if (!ex || !t) return FOHIP_E_INVALID_ARGUMENT;
    if (t->eval_strike_tile > 16 || t->fit_cashflows > FOHIP_CASHFLOWS_STOPPING_ROWS ||
        t->batch_schedule > FOHIP_BATCHES_OVERLAPPED || t->fit_schedule > FOHIP_FIT_SWEEP)
        return FOHIP_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ex->mu);
    ex->tuning = *t;
    return FOHIP_OK;
}

extern "C" int fohip_set_near_tie_window(fohip_executor *ex, double window) {
    if (!ex || !std::isfinite(window) || window < 0.0 || window > 1.0) return FOHIP_E_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(ex->mu);
    ex->near_tie_window = window;
    return FOHIP_OK;
}

// ============================================================================ group plans

namespace {

struct GroupPlan {
    std::vector<double> times, logd, q;
    std::vector<double> cash;          // per node (0 when absent)
    std::vector<uint8_t> has_cash;     // per node
    std::vector<uint32_t> obs_nodes;   // exercise nodes + expiry
    double S0 = 0.0, sigma = 0.0;
    int32_t policy = 0;
    uint32_t strike_begin = 0, strike_end = 0;
    uint64_t nt = 0, ne = 0;
    uint64_t train_seed = 0, eval_seed = 0;
    uint32_t train_domain = 0, eval_domain = 0;
    // escrowed cash dividends: the plan simulates the prepaid forward from `start` with no cash
    // events; exercise nodes observe F + prepaid[node]; chain = Greek chain factors
    bool escrowed = false;
    double start = 0.0;
    std::vector<double> prepaid, prepaid_dr;   // per node
    double chain[4] = {0.0, 1.0, 0.0, 0.0};
};

struct RunSpec {
    FitParams fp;
    uint32_t greeks = 0;
    bool store32 = false;   // binary32 spots and cashflows
    bool fused = false;     // one fused fit pass per node and the evaluation form of the rules
};

static void set_method(RunSpec &spec, int32_t precision) {
    spec.store32 = precision == FOHIP_PRECISION_F32_BULK;
    spec.fused = precision != FOHIP_PRECISION_F64;
}

struct RowOut {
    double price = 0, sd = 0, se = 0;
    bool has_sd = false;
    RowSummary sum{};
    double greek[4] = {0, 0, 0, 0}, greek_se[4] = {0, 0, 0, 0};
    uint64_t ties_training = 0, ties_evaluation = 0;   // near-tie rule applications
};

struct GroupTables {
    std::vector<DevGroup> groups;
    std::vector<double> log_drift, diffusion, dt, sqrt_dt, cash, disc0, disc_step, row_time, row_escrow, row_escrow_dr;
    std::vector<float> log_drift32, diffusion32;
    std::vector<uint32_t> node_flags, row_node;
    std::vector<StepRec> steps;
};

// Host-side discount checks (policy.rs:1081-1098): every factor used by the fit and evaluation
// must be finite and positive. Returns the offending (from, to, log discount) or false.
bool check_discount(double ld, double &d) {
    d = std::exp(ld);
    return std::isfinite(d) && d > 0.0;
}

}  // namespace

// Builds step/node/row tables of a list of groups and their HBM offsets.
static int build_tables(const std::vector<GroupPlan> &plans, const std::vector<double> &strikes, bool f32,
                        GroupTables &gt, uint64_t &path_elems, uint64_t &value_elems, uint64_t &model_count,
                        uint64_t &part_t_slots, uint64_t &part_e_slots, uint64_t &count_slots, fohip_error *err,
                        int64_t &bad_group, int64_t &bad_node) {
    (void)strikes;
    (void)f32;
    bad_group = -1;
    bad_node = -1;
    path_elems = value_elems = model_count = part_t_slots = part_e_slots = count_slots = 0;
    gt = GroupTables();
    gt.groups.reserve(plans.size());
    for (size_t gi = 0; gi < plans.size(); ++gi) {
        const GroupPlan &p = plans[gi];
        DevGroup g;
        std::memset(&g, 0, sizeof g);
        const uint32_t nodes = static_cast<uint32_t>(p.times.size());
        g.S0 = p.S0;
        g.sigma = p.sigma;
        g.escrowed = p.escrowed ? 1 : 0;
        g.start = p.escrowed ? p.start : p.S0;
        for (int k = 0; k < 4; ++k) g.chain[k] = p.escrowed ? p.chain[k] : (k == 1 ? 1.0 : 0.0);
        g.train_seed = p.train_seed;
        g.eval_seed = p.eval_seed;
        g.train_domain = p.train_domain;
        g.eval_domain = p.eval_domain;
        g.nt = static_cast<uint32_t>(p.nt);
        g.ne = static_cast<uint32_t>(p.ne);
        g.intervals = nodes - 1;
        g.obs = static_cast<uint32_t>(p.obs_nodes.size());
        g.origin_row = p.obs_nodes[0] == 0 ? 1u : 0u;
        g.fit_rows = g.obs - 1 - g.origin_row;
        g.step_base = static_cast<uint32_t>(gt.log_drift.size());
        g.node_base = static_cast<uint32_t>(gt.cash.size());
        g.row_base = static_cast<uint32_t>(gt.disc0.size());
        g.strike_begin = p.strike_begin;
        g.strike_end = p.strike_end;
        g.policy = p.policy;
        const uint32_t ns = p.strike_end - p.strike_begin;
        g.path_base = path_elems;
        path_elems += static_cast<uint64_t>(g.obs) * g.nt;
        g.value_base = value_elems;
        value_elems += static_cast<uint64_t>(ns) * g.nt;
        g.model_base = static_cast<uint32_t>(model_count);
        model_count += static_cast<uint64_t>(ns) * g.obs;
        g.count_base = static_cast<uint32_t>(count_slots);
        count_slots += static_cast<uint64_t>(ns) * g.obs;
        g.chunks_t = static_cast<uint32_t>((p.nt + kChunk - 1) / kChunk);
        g.chunks_e = static_cast<uint32_t>((p.ne + kEvalChunk - 1) / kEvalChunk);
        g.chunks_b = static_cast<uint32_t>((p.nt + kBulkChunk - 1) / kBulkChunk);
        g.part_base_t = static_cast<uint32_t>(part_t_slots);
        part_t_slots += static_cast<uint64_t>(ns) * g.chunks_t;
        g.part_base_e = static_cast<uint32_t>(part_e_slots);
        part_e_slots += static_cast<uint64_t>(ns) * g.chunks_e;
        // step.rs:27-57 transition coefficients
        const double sigma = p.sigma;
        for (uint32_t to = 1; to < nodes; ++to) {
            const uint32_t from = to - 1;
            const double dt = p.times[to] - p.times[from];
            const double interval_log_discount = p.logd[to] - p.logd[from];
            const double interval_yield = p.q[to] - p.q[from];
            const double variance = std::fma(sigma, sigma, 0.0) * dt;
            const double log_drift = -interval_log_discount - interval_yield - 0.5 * variance;
            gt.log_drift.push_back(log_drift);
            gt.diffusion.push_back(std::sqrt(variance));
            gt.log_drift32.push_back(static_cast<float>(log_drift));
            gt.diffusion32.push_back(static_cast<float>(std::sqrt(variance)));
            gt.dt.push_back(dt);
            gt.sqrt_dt.push_back(std::sqrt(dt));
        }
        const uint32_t terminal = p.obs_nodes.back();
        size_t oi = 0;
        for (uint32_t n = 0; n < nodes; ++n) {
            uint32_t flags = 0;
            while (oi < p.obs_nodes.size() && p.obs_nodes[oi] < n) ++oi;
            if (oi < p.obs_nodes.size() && p.obs_nodes[oi] == n && n > 0) flags |= kNodeObserved;
            if (n == terminal) flags |= kNodeTerminal;
            if (p.has_cash[n]) flags |= kNodeCash;
            gt.node_flags.push_back(flags);
            gt.cash.push_back(p.has_cash[n] ? p.cash[n] : 0.0);
        }
        // observation rows: discounts (policy.rs:1049-1098)
        for (uint32_t r = 0; r < g.obs; ++r) {
            const uint32_t node = p.obs_nodes[r];
            double d0 = 0.0, ds = 1.0;
            if (!check_discount(p.logd[node], d0) && bad_group < 0) {
                bad_group = static_cast<int64_t>(gi);
                bad_node = node;
            }
            if (r + 1 < g.obs) {
                const uint32_t later = p.obs_nodes[r + 1];
                if (!check_discount(p.logd[later] - p.logd[node], ds) && node != 0 && bad_group < 0) {
                    bad_group = static_cast<int64_t>(gi);
                    bad_node = node;
                }
            }
            gt.disc0.push_back(d0);
            gt.disc_step.push_back(ds);
            gt.row_time.push_back(p.times[node]);
            gt.row_node.push_back(node);
            // reconstruct_observed_stock (path.rs:844-880): the origin observes S0 and expiry
            // (the last grid node) adds no prepaid cash
            const bool adds = p.escrowed && node != 0 && node + 1 != nodes;
            gt.row_escrow.push_back(adds ? p.prepaid[node] : 0.0);
            gt.row_escrow_dr.push_back(adds ? p.prepaid_dr[node] : 0.0);
        }
        const uint32_t later_final = g.fit_rows ? p.obs_nodes[g.origin_row] : p.obs_nodes[g.obs - 1];
        double dfinal = 0.0;
        if (!check_discount(p.logd[later_final], dfinal) && bad_group < 0) {
            bad_group = static_cast<int64_t>(gi);
            bad_node = 0;
        }
        g.disc_final = dfinal;
        // the generation kernels' packed form of the same tables
        for (uint32_t to = 1, r = g.origin_row; to < nodes; ++to) {
            StepRec rec;
            std::memset(&rec, 0, sizeof rec);
            const size_t si = g.step_base + to - 1, ni = g.node_base + to;
            rec.log_drift = gt.log_drift[si];
            rec.diffusion = gt.diffusion[si];
            rec.dt = gt.dt[si];
            rec.sqrt_dt = gt.sqrt_dt[si];
            rec.cash = gt.cash[ni];
            rec.flags = gt.node_flags[ni];
            if (rec.flags & kNodeObserved) {
                rec.disc0 = gt.disc0[g.row_base + r];
                rec.row_time = gt.row_time[g.row_base + r];
                ++r;
            }
            gt.steps.push_back(rec);
        }
        gt.groups.push_back(g);
    }
    (void)err;
    return FOHIP_OK;
}

namespace {

// Rust Display of f64 values in messages uses the shortest round-trip form; %.17g is close enough
// for diagnostics and is exact.
const char *regression_message(int code, char *buf, size_t cap, int degree, const RowError &e) {
    switch (code) {
        case FOHIP_E_INSUFFICIENT_OBSERVATIONS:
            std::snprintf(buf, cap, "regression has %lld observations but degree %d requires at least %d",
                          (long long)e.index, degree, degree + 1);
            break;
        case FOHIP_E_DEGENERATE_PREDICTOR:
            std::snprintf(buf, cap, "degree %d regression predictor scale is degenerate; observed %.17g", degree,
                          e.observed);
            break;
        case FOHIP_E_RANK_DEFICIENT:
            std::snprintf(buf, cap,
                          "degree %d polynomial design has numerical rank %lld of %d at rank tolerance %.17g",
                          degree, (long long)e.index, degree + 1, e.limit);
            break;
        case FOHIP_E_CONDITION_LIMIT:
            std::snprintf(buf, cap,
                          "degree %d polynomial design condition number %.17g exceeds configured limit %.17g",
                          degree, e.observed, e.limit);
            break;
        case FOHIP_E_JACOBI_NOT_CONVERGED:
            std::snprintf(buf, cap, "polynomial condition diagnostic did not converge for %lld columns after 32 sweeps",
                          (long long)e.index);
            break;
        default:
            std::snprintf(buf, cap, "polynomial regression produced non-finite arithmetic at basis degree %lld",
                          (long long)e.index);
    }
    return buf;
}

}  // namespace

namespace {

// Device buffers of one batch. Grown mode (the default) claims the executor's own buffers, growing
// them with hipMalloc as needed, and stages transfers through its page-locked buffer. Fixed mode
// carves every buffer out of one caller-provided device region (the allocation-free entry point):
// claims only advance an offset, `overflow` records a request larger than the region, and
// transfers go through pageable host memory, so no device or pinned allocation happens.
struct Workspace {
    fohip_executor *ex = nullptr;
    char *region = nullptr;
    size_t capacity = 0, used = 0, needed = 0;
    bool overflow = false;
    std::vector<char> pageable;
    explicit Workspace(fohip_executor *e) : ex(e) {}
    Workspace(fohip_executor *e, void *base, size_t bytes) : ex(e), region(static_cast<char *>(base)), capacity(bytes) {}
    void begin_batch() {
        used = needed = 0;
        overflow = false;
    }
    hipError_t claim(DevBuf &own, size_t bytes, void **ptr) {
        if (!region) {
            hipError_t e = own.ensure(bytes);
            *ptr = own.ptr;
            return e;
        }
        const size_t off = (used + 255) & ~size_t(255);
        needed = off + bytes;
        if (needed > capacity) {
            overflow = true;
            *ptr = nullptr;
            return hipSuccess;
        }
        *ptr = region + off;
        used = needed;
        return hipSuccess;
    }
    hipError_t host_staging(PinnedBuf &own, size_t bytes, char **ptr) {
        if (!region) {
            hipError_t e = own.ensure(bytes);
            *ptr = own.ptr;
            return e;
        }
        if (pageable.size() < bytes) pageable.resize(bytes);
        *ptr = pageable.data();
        return hipSuccess;
    }
};

}  // namespace

namespace {

// ---- fit sweep schedule
//
// A group of c fused-fit chunks is swept by a team of m = ceil(c / chunks_per_block) blocks, each
// holding at most chunks_per_block consecutive chunks (the counts differ by at most one). The
// blocks of the launch take the entries by ticket: teams in order of decreasing cost
// (fit_rows x (chunks per member + the step's gathering latency), so the longest sweeps start
// first and the short ones fill in at the end), each team's members on consecutive tickets.
constexpr double kSweepStepLatency = 4.0;   // a step's gathering and solve, in chunk passes
constexpr uint32_t kSweepSpinLimit = 1u << 22;   // polls of one wait (each a device-memory round trip): seconds

struct SweepSchedule {
    std::vector<SweepEntry> entries;   // ticket order
    uint32_t teams = 0;
    uint32_t most_members = 0;
    uint64_t records = 0;              // partial records: 2 per chunk of every team
};

void plan_sweep(const uint32_t *chunks, const uint32_t *fit_rows, uint32_t groups, uint32_t chunks_per_block,
                SweepSchedule &out) {
    out = SweepSchedule();
    struct Team {
        uint32_t group, members;
        double cost;
    };
    std::vector<Team> teams;
    for (uint32_t g = 0; g < groups; ++g) {
        if (fit_rows[g] == 0 || chunks[g] == 0) continue;
        const uint32_t m = (chunks[g] + chunks_per_block - 1) / chunks_per_block;
        const uint32_t per = (chunks[g] + m - 1) / m;
        teams.push_back({g, m, static_cast<double>(fit_rows[g]) * (per + kSweepStepLatency)});
    }
    std::stable_sort(teams.begin(), teams.end(), [](const Team &a, const Team &b) { return a.cost > b.cost; });
    for (uint32_t id = 0; id < teams.size(); ++id) {
        const Team &T = teams[id];
        const uint32_t c = chunks[T.group];
        for (uint32_t i = 0; i < T.members; ++i) {
            SweepEntry e{};
            e.group = T.group;
            e.chunk_begin = static_cast<uint32_t>(static_cast<uint64_t>(c) * i / T.members);
            e.chunk_end = static_cast<uint32_t>(static_cast<uint64_t>(c) * (i + 1) / T.members);
            e.members = T.members;
            e.team = id;
            e.leader = i == 0 ? 1u : 0u;
            e.part_base = out.records;
            out.entries.push_back(e);
        }
        out.records += 2ull * c;
        out.most_members = std::max(out.most_members, T.members);
    }
    out.teams = static_cast<uint32_t>(teams.size());
}

}  // namespace

extern "C" int fohip_sweep_schedule(const uint32_t *chunks, const uint32_t *fit_rows, uint32_t groups,
                                    uint32_t chunks_per_block, uint32_t *entries, uint64_t entry_capacity,
                                    uint64_t *entry_count, uint32_t *teams) {
    if ((groups && (!chunks || !fit_rows)) || !entry_count || !teams || chunks_per_block == 0)
        return FOHIP_E_INVALID_ARGUMENT;
    SweepSchedule plan;
    plan_sweep(chunks, fit_rows, groups, chunks_per_block, plan);
    *entry_count = plan.entries.size();
    *teams = plan.teams;
    if (plan.entries.size() > entry_capacity || (!entries && !plan.entries.empty())) return FOHIP_E_INVALID_ARGUMENT;
    for (size_t e = 0; e < plan.entries.size(); ++e) {
        const SweepEntry &x = plan.entries[e];
        uint32_t *w = entries + 8 * e;
        w[0] = x.group;
        w[1] = x.chunk_begin;
        w[2] = x.chunk_end;
        w[3] = x.members;
        w[4] = x.team;
        w[5] = x.leader;
        w[6] = static_cast<uint32_t>(x.part_base);
        w[7] = static_cast<uint32_t>(x.part_base >> 32);
    }
    return FOHIP_OK;
}

namespace {

// One memory-bounded batch of consecutive groups on its way through the device: the host tables
// built for it, the slot buffers it claimed and the host copies its results come back into.
struct BatchRun {
    int slot = 0;   // device buffers
    int io = 0;     // events and page-locked staging
    std::vector<GroupPlan> plans;
    uint32_t row0 = 0, nrows = 0, ngroups = 0;
    std::vector<double> bstrikes, binv;
    GroupTables gt;
    uint64_t path_elems = 0, value_elems = 0, model_count = 0, part_t = 0, part_e = 0, count_slots = 0;
    int64_t bad_group = -1, bad_node = -1;
    uint64_t znorm_elems = 0;
    std::vector<uint32_t> strike_group;
    uint32_t max_strikes = 0, max_obs = 0, max_fit = 0;
    std::vector<uint2> items_t, items_e, items_b;
    std::vector<uint32_t> rows_sorted, step_items, step_rows, step_items_b;
    static constexpr int kTableParts = 24;
    size_t part_bytes[kTableParts] = {};
    const void *part_host[kTableParts] = {};
    size_t part_off[kTableParts] = {};
    size_t tables_bytes = 0;
    bool stored_normals = false, escrow = false;
    // device side
    void *paths_d = nullptr, *values_d = nullptr, *stops_d = nullptr, *models_d = nullptr, *part_t_d = nullptr,
         *part_e_d = nullptr, *counts_d = nullptr, *nosup_d = nullptr, *errors_d = nullptr, *fail_t_d = nullptr,
         *fail_e_d = nullptr, *eval_out_d = nullptr, *summary_d = nullptr, *ties_d = nullptr, *tables_d = nullptr,
         *znorm_d = nullptr, *replay_d = nullptr;
    DevTables t = {};
    const uint2 *it = nullptr, *ite = nullptr, *itb = nullptr;
    const uint32_t *rows = nullptr;
    // fit sweep (single-strike groups): schedule, device state and the abort word read back
    bool sweep = false;
    uint32_t sweep_slots = 0;
    SweepSchedule sw;
    const SweepEntry *sweep_entries_d = nullptr;
    void *sweep_sync_d = nullptr, *sweep_part_d = nullptr, *sweep_abort_d = nullptr;
    uint32_t sweep_abort = 0;
    char *staging = nullptr;
    // host copies of the results
    std::vector<unsigned long long> fail_t, fail_e, ties_h;
    std::vector<RowError> row_err;
    std::vector<double> eval_out;
    std::vector<RowSummary> summary;
    std::vector<Model> models_h;
    std::vector<uint32_t> counts_h;
    struct Download {
        void *host;
        void *const *dev;   // read when the transfer is queued, after the claims
        size_t bytes;
    };
    std::vector<Download> downloads() {
        return {{fail_t.data(), (&fail_t_d), ngroups * sizeof(unsigned long long)},
                {fail_e.data(), (&fail_e_d), ngroups * sizeof(unsigned long long)},
                {row_err.data(), (&errors_d), nrows * sizeof(RowError)},
                {eval_out.data(), (&eval_out_d), eval_out.size() * sizeof(double)},
                {summary.data(), (&summary_d), nrows * sizeof(RowSummary)},
                {models_h.data(), (&models_d), models_h.size() * sizeof(Model)},
                {counts_h.data(), (&counts_d), counts_h.size() * sizeof(uint32_t)},
                {ties_h.data(), (&ties_d), ties_h.size() * sizeof(unsigned long long)},
                {&sweep_abort, (&sweep_abort_d), sweep ? sizeof(uint32_t) : 0}};
    }
};

// Waits for both streams of the executor when a request returns, so that no batch of it is still
// in flight when the next request reuses the slots.
struct DrainOnExit {
    fohip_executor *ex;
    ~DrainOnExit() {
        (void)hipStreamSynchronize(ex->gen);
        (void)hipStreamSynchronize(ex->stream);
    }
};

}  // namespace

// Runs the full pipeline for a list of groups, in memory-bounded batches of consecutive groups.
//
// Serial schedule: each batch's simulation, backward induction and evaluation run in turn, while
// the host builds the next batch's tables.
// Overlapped schedule: batch b's backward induction runs on the executor's stream while batch
// b + 1's simulation and batch b - 1's evaluation run on the generation stream. The host builds
// batch b + 1's tables while the device works, finishes batch b - 1 (its results and failures,
// in request order) before reusing its slot for b + 1, and queues b + 1's simulation ahead of b's
// evaluation so that the generation stream never waits behind a fit it does not need. Results
// do not depend on the schedule: every reduction tree is fixed per group.
static int run_groups(fohip_executor *ex, const std::vector<GroupPlan> &all, const std::vector<double> &strikes,
                      const RunSpec &spec, std::vector<RowOut> &out, std::vector<Model> *models_out,
                      std::vector<uint32_t> *counts_out, fohip_error *err, Workspace *fixed = nullptr) {
    Workspace grown(ex);
    Workspace &ws = fixed ? *fixed : grown;
    HIP_TRY(hipSetDevice(ex->device));
    DrainOnExit drain{ex};
    hipStream_t fit_st = ex->stream, gen_st = ex->gen;
    // the fixed region of the allocation-free entry point holds one batch at a time
    const bool overlapped = !fixed && ex->tuning.batch_schedule == FOHIP_BATCHES_OVERLAPPED;
    const bool f32 = spec.store32;
    const size_t real = f32 ? sizeof(float) : sizeof(double);
    out.assign(strikes.size(), RowOut());
    // Cashflow encoding of the fused backward induction, one choice per request: stopping rows
    // (1 byte per strike and path) when every group has at most 255 observation rows and the
    // tuning asks for them. The default keeps value columns: the stopping rows' gathers cost more
    // than the bytes they save (measured on the 64-strike, 256-strike and single-strike shapes).
    bool use_stops = false;
    uint32_t most_strikes = 0;
    size_t most_obs = 0;
    for (const GroupPlan &p : all) {
        most_strikes = std::max<uint32_t>(most_strikes, p.strike_end - p.strike_begin);
        most_obs = std::max<size_t>(most_obs, p.obs_nodes.size());
    }
    if (spec.fused) {
        const uint32_t mode = ex->tuning.fit_cashflows;
        use_stops = most_obs <= 255 && mode == FOHIP_CASHFLOWS_STOPPING_ROWS;
    }
    // The fit sweep, one choice per request and only when the tuning asks for it: fused method,
    // value-column cashflows, one strike per group, at most 2^28 training paths per group (one
    // buffer resource's byte range), and no team larger than the device's CU count (every team's
    // blocks can run at once: one per CU). Per-node launches stay the default: on the ragged batch
    // the sweep's fit takes 1.6 s against 0.54 s, because each backward step of a team waits about
    // 16 us for its members' records and its solve with one block per CU and nothing to overlap.
    uint32_t sweep_slots = 0;
    const uint32_t sweep_chunks = f32 ? kSweepChunks32 : kSweepChunks64;
    if (spec.fused && !use_stops && most_strikes == 1 && ex->tuning.fit_schedule == FOHIP_FIT_SWEEP) {
        const int win = ex->near_tie_window != 0.0 ? 1 : 0;
        int64_t &slots = ex->sweep_slots[f32 ? 1 : 0][win];
        if (slots < 0) {
            int cus = 0;
            const int per_cu = sweep_resident_blocks(f32, win != 0);
            if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ex->device) != hipSuccess) cus = 0;
            (void)hipGetLastError();
            slots = static_cast<int64_t>(per_cu) * cus;
        }
        int cus = 0;
        if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ex->device) != hipSuccess) cus = 0;
        bool fits = slots > 0 && cus > 0;
        for (const GroupPlan &p : all) {
            const uint64_t cb = (p.nt + kBulkChunk - 1) / kBulkChunk;
            if (p.nt >= (1ull << 28) || (cb + sweep_chunks - 1) / sweep_chunks > static_cast<uint64_t>(cus))
                fits = false;
        }
        if (fits) sweep_slots = static_cast<uint32_t>(slots);
    }
    // The evaluation's strike tile, one per request (results do not depend on it); windowed
    // near-tie counts run the evaluation's diagnostic form over single-strike tiles.
    const bool greeks = spec.greeks != 0;
    const uint32_t tile = ex->near_tie_window != 0.0 ? 1u
                                                     : eval_tile_for(most_strikes, greeks, ex->tuning.eval_strike_tile);
    // Batches of consecutive groups bounded by memory (paths, cashflows and stored normals within
    // the slot budget) and by the launch grids: a dispatch holds at most 2^32 - 1 work-items, so
    // no kernel of a batch may have more than kMaxBlocks blocks of 256 threads. A grid is (work
    // items of the batch, rounded up to 8 for the XCD mapping) x (strike columns of the batch).
    // An explicit batch_mib gives both schedules the same batches.
    constexpr uint64_t kMaxBlocks = 0xFFFFFFFFull / 256;   // 256-thread blocks
    const size_t budget = ex->tuning.batch_mib ? static_cast<size_t>(ex->tuning.batch_mib) << 20
                                               : overlapped ? ex->batch_budget / 2 : ex->batch_budget;
    auto grid_blocks = [&](uint64_t items_t, uint64_t items_e, uint64_t items_b, uint64_t ns) {
        const uint64_t fit = ((items_t + 7) / 8) * 8 * ns;
        const uint64_t eval = ((items_e + 7) / 8) * 8 * ((ns + tile - 1) / tile);
        const uint64_t bulk = ((items_b + 7) / 8) * 8 * ((ns + kBulkStrikes - 1) / kBulkStrikes);
        return std::max({fit, eval, bulk, items_e});
    };
    std::vector<std::pair<size_t, size_t>> batches;
    {
        size_t begin = 0, bytes = 0;
        uint64_t it_t = 0, it_e = 0, it_b = 0, cols = 0;
        for (size_t g = 0; g < all.size(); ++g) {
            const GroupPlan &p = all[g];
            const uint64_t ns = p.strike_end - p.strike_begin;
            size_t need = p.obs_nodes.size() * p.nt * real + ns * p.nt * (use_stops ? 1 : real);
            if (spec.fused && ns > 1)   // stored evaluation normals
                need += (p.times.size() - 1) * p.ne * sizeof(double);
            const uint64_t ct = (p.nt + kChunk - 1) / kChunk, ce = (p.ne + kEvalChunk - 1) / kEvalChunk,
                           cb = (p.nt + kBulkChunk - 1) / kBulkChunk;
            if (grid_blocks(ct, ce, cb, ns) > kMaxBlocks)
                return fail(err, FOHIP_E_UNSUPPORTED, 0,
                            "group %zu (%llu strikes, %llu + %llu paths) exceeds one launch grid", g,
                            (unsigned long long)ns, (unsigned long long)p.nt, (unsigned long long)p.ne);
            if (g > begin && (bytes + need > budget ||
                              grid_blocks(it_t + ct, it_e + ce, it_b + cb, std::max(cols, ns)) > kMaxBlocks)) {
                batches.push_back({begin, g});
                begin = g;
                bytes = 0;
                it_t = it_e = it_b = cols = 0;
            }
            bytes += need;
            it_t += ct;
            it_e += ce;
            it_b += cb;
            cols = std::max(cols, ns);
        }
        batches.push_back({begin, all.size()});
    }
    if (models_out) models_out->clear();
    if (counts_out) counts_out->clear();
    ex->last_models.clear();

    // ---- host side of a batch: plans rebased to the batch, tables, work items, result vectors
    auto prepare = [&](size_t b) {
        auto B = std::make_unique<BatchRun>();
        BatchRun &R = *B;
        R.slot = overlapped ? static_cast<int>(b & 1) : 0;
        R.io = fixed ? 0 : static_cast<int>(b & 1);
        const auto [gb, ge] = batches[b];
        R.plans.assign(all.begin() + gb, all.begin() + ge);
        // strike rows of this batch are contiguous: rebase to the batch
        R.row0 = R.plans.front().strike_begin;
        const uint32_t row1 = R.plans.back().strike_end;
        R.nrows = row1 - R.row0;
        R.bstrikes.assign(strikes.begin() + R.row0, strikes.begin() + row1);
        R.binv.resize(R.bstrikes.size());   // 1/K, correctly rounded as on the device
        for (size_t i = 0; i < R.bstrikes.size(); ++i) R.binv[i] = 1.0 / R.bstrikes[i];
        for (auto &p : R.plans) {
            p.strike_begin -= R.row0;
            p.strike_end -= R.row0;
        }
        build_tables(R.plans, R.bstrikes, f32, R.gt, R.path_elems, R.value_elems, R.model_count, R.part_t, R.part_e,
                     R.count_slots, err, R.bad_group, R.bad_node);
        for (DevGroup &dg : R.gt.groups) {   // evaluation normals, [intervals][ne] per group
            dg.znorm_base = R.znorm_elems;
            R.znorm_elems += static_cast<uint64_t>(dg.intervals) * dg.ne;
        }
        R.ngroups = static_cast<uint32_t>(R.plans.size());
        R.strike_group.resize(R.nrows);
        for (uint32_t g = 0; g < R.ngroups; ++g) {
            for (uint32_t s = R.plans[g].strike_begin; s < R.plans[g].strike_end; ++s) R.strike_group[s] = g;
            R.max_strikes = std::max(R.max_strikes, R.plans[g].strike_end - R.plans[g].strike_begin);
            R.max_obs = std::max(R.max_obs, R.gt.groups[g].obs);
            R.max_fit = std::max(R.max_fit, R.gt.groups[g].fit_rows);
        }
        // work items ordered by fit rows (descending) so each backward step uses a prefix
        std::vector<uint32_t> order(R.ngroups);
        for (uint32_t g = 0; g < R.ngroups; ++g) order[g] = g;
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t c) { return R.gt.groups[a].fit_rows > R.gt.groups[c].fit_rows; });
        R.step_items.assign(R.max_fit + 1, 0);
        R.step_rows.assign(R.max_fit + 1, 0);
        R.step_items_b.assign(R.max_fit + 1, 0);
        for (uint32_t g : order) {
            for (uint32_t c = 0; c < R.gt.groups[g].chunks_t; ++c) R.items_t.push_back(make_uint2(g, c));
            for (uint32_t c = 0; c < R.gt.groups[g].chunks_b; ++c) R.items_b.push_back(make_uint2(g, c));
            for (uint32_t s = R.plans[g].strike_begin; s < R.plans[g].strike_end; ++s) R.rows_sorted.push_back(s);
        }
        for (uint32_t g = 0; g < R.ngroups; ++g)
            for (uint32_t c = 0; c < R.gt.groups[g].chunks_e; ++c) R.items_e.push_back(make_uint2(g, c));
        for (uint32_t j = 0; j <= R.max_fit; ++j) {
            uint32_t ni = 0, nr = 0, nb = 0;
            for (uint32_t g : order) {
                if (R.gt.groups[g].fit_rows > j || j == R.max_fit) {
                    ni += R.gt.groups[g].chunks_t;
                    nb += R.gt.groups[g].chunks_b;
                    nr += R.plans[g].strike_end - R.plans[g].strike_begin;
                }
            }
            R.step_items[j] = ni;
            R.step_items_b[j] = nb;
            R.step_rows[j] = nr;
        }
        if (sweep_slots) {
            std::vector<uint32_t> ch(R.ngroups), fr(R.ngroups);
            for (uint32_t g = 0; g < R.ngroups; ++g) {
                ch[g] = R.gt.groups[g].chunks_b;
                fr[g] = R.gt.groups[g].fit_rows;
            }
            plan_sweep(ch.data(), fr.data(), R.ngroups, sweep_chunks, R.sw);
            R.sweep = !R.sw.entries.empty();
            R.sweep_slots = sweep_slots;
        }
        const GroupTables &gt = R.gt;
        const size_t bytes[BatchRun::kTableParts] = {
            gt.groups.size() * sizeof(DevGroup), gt.log_drift.size() * 8, gt.diffusion.size() * 8,
            gt.log_drift32.size() * 4, gt.diffusion32.size() * 4, gt.dt.size() * 8, gt.sqrt_dt.size() * 8,
            gt.cash.size() * 8, gt.node_flags.size() * 4, gt.disc0.size() * 8, gt.disc_step.size() * 8,
            gt.row_time.size() * 8, gt.row_node.size() * 4, R.bstrikes.size() * 8, R.strike_group.size() * 4,
            gt.row_escrow.size() * 8, gt.row_escrow_dr.size() * 8, R.binv.size() * 8,
            R.items_t.size() * sizeof(uint2), R.items_e.size() * sizeof(uint2), R.items_b.size() * sizeof(uint2),
            R.rows_sorted.size() * sizeof(uint32_t), gt.steps.size() * sizeof(StepRec),
            R.sweep ? R.sw.entries.size() * sizeof(SweepEntry) : 0};
        const void *host[BatchRun::kTableParts] = {
            gt.groups.data(), gt.log_drift.data(), gt.diffusion.data(), gt.log_drift32.data(),
            gt.diffusion32.data(), gt.dt.data(), gt.sqrt_dt.data(), gt.cash.data(),
            gt.node_flags.data(), gt.disc0.data(), gt.disc_step.data(), gt.row_time.data(),
            gt.row_node.data(), R.bstrikes.data(), R.strike_group.data(),
            gt.row_escrow.data(), gt.row_escrow_dr.data(), R.binv.data(),
            R.items_t.data(), R.items_e.data(), R.items_b.data(), R.rows_sorted.data(), gt.steps.data(),
            R.sw.entries.data()};
        for (int i = 0; i < BatchRun::kTableParts; ++i) {
            R.part_bytes[i] = bytes[i];
            R.part_host[i] = host[i];
            R.part_off[i] = R.tables_bytes;
            R.tables_bytes += (bytes[i] + 255) & ~size_t(255);
        }
        // Several strike tiles read one set of evaluation normals from HBM instead of each
        // regenerating Philox and ZIGNOR (the normals are bitwise those of the streams). Single-
        // strike tiles always regenerate: that form has no stored-normal kernel.
        R.stored_normals = spec.fused && tile > 1 && (R.max_strikes + tile - 1) / tile >= 2 && R.znorm_elems > 0;
        for (const GroupPlan &P : R.plans) R.escrow = R.escrow || P.escrowed;
        R.fail_t.resize(R.ngroups);
        R.fail_e.resize(R.ngroups);
        R.row_err.resize(R.nrows);
        R.eval_out.resize(static_cast<size_t>(R.nrows) * 16);
        R.summary.resize(R.nrows);
        R.models_h.resize(models_out ? R.model_count : 0);
        R.counts_h.resize(counts_out ? R.count_slots : 0);
        // tied training decisions counted with atomics by the F64 fit passes and the last fitted
        // row; the fused pass's are in the row summary and the evaluation's in eval_out
        R.ties_h.resize(R.nrows);
        return B;
    };

    // ---- slot buffers, upload and simulation (generation stream); backward induction (fit stream)
    auto issue_front = [&](BatchRun &R) -> int {
        BatchSlot &S = ex->slot[R.slot];
        BatchSlot &E = ex->slot[R.io];
        // every buffer of the batch is claimed (and a fixed workspace found large enough) before
        // any work is queued
        ws.begin_batch();
        HIP_TRY(ws.claim(S.tables, R.tables_bytes, &R.tables_d));
        HIP_TRY(ws.claim(S.paths, std::max<uint64_t>(R.path_elems, 1) * real, &R.paths_d));
        if (use_stops)
            HIP_TRY(ws.claim(S.stops, (std::max<uint64_t>(R.value_elems, 1) + 3) & ~uint64_t(3), &R.stops_d));
        else
            HIP_TRY(ws.claim(S.values, std::max<uint64_t>(R.value_elems, 1) * real, &R.values_d));
        HIP_TRY(ws.claim(S.models, std::max<uint64_t>(R.model_count, 1) * sizeof(Model), &R.models_d));
        HIP_TRY(ws.claim(S.part_t, std::max<uint64_t>(R.part_t, 1) * kPartialStride * sizeof(double), &R.part_t_d));
        HIP_TRY(ws.claim(S.part_e, std::max<uint64_t>(R.part_e, 1) * kPartialStride * sizeof(double), &R.part_e_d));
        HIP_TRY(ws.claim(S.counts, std::max<uint64_t>(R.count_slots, 1) * sizeof(uint32_t), &R.counts_d));
        HIP_TRY(ws.claim(S.nosup, std::max<uint64_t>(R.count_slots, 1) * sizeof(uint32_t), &R.nosup_d));
        HIP_TRY(ws.claim(S.errors, std::max<uint32_t>(R.nrows, 1) * sizeof(RowError), &R.errors_d));
        HIP_TRY(ws.claim(S.fail_t, R.ngroups * sizeof(unsigned long long), &R.fail_t_d));
        HIP_TRY(ws.claim(S.fail_e, R.ngroups * sizeof(unsigned long long), &R.fail_e_d));
        HIP_TRY(ws.claim(S.eval_out, std::max<uint32_t>(R.nrows, 1) * 16 * sizeof(double), &R.eval_out_d));
        HIP_TRY(ws.claim(S.summary, std::max<uint32_t>(R.nrows, 1) * sizeof(RowSummary), &R.summary_d));
        HIP_TRY(ws.claim(S.ties, std::max<uint32_t>(R.nrows, 1) * sizeof(unsigned long long), &R.ties_d));
        HIP_TRY(ws.claim(S.replay, 4 * sizeof(double), &R.replay_d));
        if (R.stored_normals) HIP_TRY(ws.claim(S.znorm, R.znorm_elems * sizeof(double), &R.znorm_d));
        if (R.sweep) {
            HIP_TRY(ws.claim(S.sweep_sync, (static_cast<size_t>(R.sw.teams) + 2) * sizeof(uint32_t), &R.sweep_sync_d));
            HIP_TRY(ws.claim(S.sweep_part, std::max<uint64_t>(R.sw.records, 1) * kSweepStride * sizeof(double),
                             &R.sweep_part_d));
            if (R.sweep_sync_d) R.sweep_abort_d = static_cast<uint32_t *>(R.sweep_sync_d) + R.sw.teams;
        }
        if (ws.overflow) {
            return fail(err, FOHIP_E_UNSUPPORTED, 0,
                        "request needs %llu bytes of device scratch; the allocation-free entry point holds %llu",
                        (unsigned long long)ws.needed, (unsigned long long)ws.capacity);
        }
        // one upload of every table, item list and row list; results come back through the
        // staging area too, after the tables
        size_t staged = R.tables_bytes;
        for (const auto &d : R.downloads()) staged += (d.bytes + 255) & ~size_t(255);
        HIP_TRY(ws.host_staging(E.staging, staged, &R.staging));
        for (int i = 0; i < BatchRun::kTableParts; ++i)
            if (R.part_bytes[i]) std::memcpy(R.staging + R.part_off[i], R.part_host[i], R.part_bytes[i]);
        char *tb = static_cast<char *>(R.tables_d);
        HIP_TRY(hipMemcpyAsync(tb, R.staging, R.tables_bytes, hipMemcpyHostToDevice, gen_st));
        DevTables &t = R.t;
        const size_t *o = R.part_off;
        t.groups = reinterpret_cast<const DevGroup *>(tb + o[0]);
        t.log_drift = reinterpret_cast<const double *>(tb + o[1]);
        t.diffusion = reinterpret_cast<const double *>(tb + o[2]);
        t.log_drift32 = reinterpret_cast<const float *>(tb + o[3]);
        t.diffusion32 = reinterpret_cast<const float *>(tb + o[4]);
        t.dt = reinterpret_cast<const double *>(tb + o[5]);
        t.sqrt_dt = reinterpret_cast<const double *>(tb + o[6]);
        t.cash = reinterpret_cast<const double *>(tb + o[7]);
        t.node_flags = reinterpret_cast<const uint32_t *>(tb + o[8]);
        t.disc0 = reinterpret_cast<const double *>(tb + o[9]);
        t.disc_step = reinterpret_cast<const double *>(tb + o[10]);
        t.row_time = reinterpret_cast<const double *>(tb + o[11]);
        t.row_node = reinterpret_cast<const uint32_t *>(tb + o[12]);
        t.strikes = reinterpret_cast<const double *>(tb + o[13]);
        t.strike_group = reinterpret_cast<const uint32_t *>(tb + o[14]);
        t.row_escrow = reinterpret_cast<const double *>(tb + o[15]);
        t.row_escrow_dr = reinterpret_cast<const double *>(tb + o[16]);
        t.inv_strikes = reinterpret_cast<const double *>(tb + o[17]);
        R.it = reinterpret_cast<const uint2 *>(tb + o[18]);
        R.ite = reinterpret_cast<const uint2 *>(tb + o[19]);
        R.itb = reinterpret_cast<const uint2 *>(tb + o[20]);
        R.rows = reinterpret_cast<const uint32_t *>(tb + o[21]);
        t.steps = reinterpret_cast<const StepRec *>(tb + o[22]);
        R.sweep_entries_d = reinterpret_cast<const SweepEntry *>(tb + o[23]);
        // rules zeroed, exercise and no-support counters zeroed, failure keys set to "none"
        static_assert(sizeof(Model) % 4 == 0 && sizeof(RowError) % 4 == 0, "fills write 32-bit words");
        FillList fills = {};
        auto add_fill = [&fills](void *ptr, size_t nbytes, uint32_t value) {
            fills.region[fills.count++] = FillRegion{static_cast<uint32_t *>(ptr), nbytes / 4, value, 0};
        };
        add_fill(R.models_d, R.model_count * sizeof(Model), 0u);
        add_fill(R.counts_d, R.count_slots * sizeof(uint32_t), 0u);
        add_fill(R.nosup_d, R.count_slots * sizeof(uint32_t), 0u);
        add_fill(R.errors_d, R.nrows * sizeof(RowError), 0xFFFFFFFFu);
        add_fill(R.fail_t_d, R.ngroups * sizeof(unsigned long long), 0xFFFFFFFFu);
        add_fill(R.fail_e_d, R.ngroups * sizeof(unsigned long long), 0xFFFFFFFFu);
        add_fill(R.ties_d, R.ties_h.size() * sizeof(unsigned long long), 0u);
        // every path starts stopped at expiry: stopping byte 0
        if (use_stops) add_fill(R.stops_d, (R.value_elems + 3) & ~uint64_t(3), 0u);
        // team arrival counters, the abort word and the ticket counter of the fit sweep
        if (R.sweep) add_fill(R.sweep_sync_d, (static_cast<size_t>(R.sw.teams) + 2) * sizeof(uint32_t), 0u);
        HIP_TRY(launch_fill(gen_st, fills));
        // ---- training simulation
        PROF_LAUNCH(gen_st, "k_simulate",
                    launch_simulate(gen_st, f32, t, R.it, static_cast<uint32_t>(R.items_t.size()), R.paths_d, 0,
                                    static_cast<unsigned long long *>(R.fail_t_d)));
        HIP_TRY(hipEventRecord(E.simulated, gen_st));
        HIP_TRY(hipStreamWaitEvent(fit_st, E.simulated, 0));
        // ---- backward induction
        uint8_t *stops = use_stops ? static_cast<uint8_t *>(R.stops_d) : nullptr;
        unsigned long long *ties_t = static_cast<unsigned long long *>(R.ties_d);
        Model *models = static_cast<Model *>(R.models_d);
        RowError *errors = static_cast<RowError *>(R.errors_d);
        double *part_t = static_cast<double *>(R.part_t_d);
        FitParams fp = spec.fp;
        fp.tie_window = ex->near_tie_window;
        if (R.sweep) {   // every backward step of the batch in one persistent launch
            SweepPlan plan{R.sweep_entries_d, static_cast<uint32_t *>(R.sweep_sync_d),
                           static_cast<double *>(R.sweep_part_d), static_cast<uint32_t>(R.sw.entries.size()),
                           R.sw.teams, kSweepSpinLimit, 0u};
            PROF_LAUNCH(fit_st, "k_fit_sweep",
                        launch_fit_sweep(fit_st, f32, t, fp, plan, R.sweep_slots, R.paths_d, R.values_d, models,
                                         errors));
        }
        for (uint32_t j = 0; j < (R.sweep ? 0u : R.max_fit); ++j) {
            const uint32_t ni = R.step_items[j], nr = R.step_rows[j];
            if (spec.fused) {   // one fused HBM pass per backward step
                PROF_LAUNCH(fit_st, "k_fit_bulk_step",
                            launch_fit_bulk_step(fit_st, f32, t, fp, j, R.itb, R.step_items_b[j], R.max_strikes,
                                                 R.paths_d, R.values_d, stops, models, part_t, errors));
                PROF_LAUNCH(fit_st, "k_fit_bulk_reduce",
                            launch_fit_bulk_reduce(fit_st, t, fp, j, R.rows, nr, part_t, models, errors));
                continue;
            }
            PROF_LAUNCH(fit_st, "k_fit_prepare",
                        launch_fit_prepare(fit_st, t, fp, j, R.it, ni, R.max_strikes, R.paths_d, R.values_d, models,
                                           part_t, errors, ties_t));
            PROF_LAUNCH(fit_st, "k_fit_reduce_welford",
                        launch_fit_reduce_welford(fit_st, t, fp, j, R.rows, nr, part_t, models, errors));
            for (int k = 0; k <= fp.degree; ++k) {
                PROF_LAUNCH(fit_st, "k_fit_moments",
                            launch_fit_moments(fit_st, t, fp, j, k, R.it, ni, R.max_strikes, R.paths_d, R.values_d,
                                               models, part_t, errors));
                PROF_LAUNCH(fit_st, "k_fit_reduce_moments",
                            launch_fit_reduce_moments(fit_st, t, fp, j, k, R.rows, nr, part_t, models, errors));
            }
        }
        if (spec.fused)   // rank/condition gates of every rule of the backward pass at once
            PROF_LAUNCH(fit_st, "k_fit_gate",
                        launch_fit_gate(fit_st, t, fp, R.rows, R.nrows, models,
                                        static_cast<uint32_t>(R.model_count), errors));
        PROF_LAUNCH(fit_st, "k_fit_final",
                    launch_fit_final(fit_st, f32, spec.fused, t, fp, R.it, static_cast<uint32_t>(R.items_t.size()),
                                     R.max_strikes, R.paths_d, R.values_d, stops, models, part_t, errors, ties_t));
        PROF_LAUNCH(fit_st, "k_fit_reduce_origin",
                    launch_fit_reduce_origin(fit_st, t, fp, R.rows, R.nrows, part_t, models, errors));
        HIP_TRY(hipEventRecord(E.fitted, fit_st));
        return FOHIP_OK;
    };

    // ---- evaluation and downloads (generation stream, after the batch's backward induction)
    auto issue_back = [&](BatchRun &R) -> int {
        BatchSlot &E = ex->slot[R.io];
        const DevTables &t = R.t;
        FitParams fp = spec.fp;
        fp.tie_window = ex->near_tie_window;
        Model *models = static_cast<Model *>(R.models_d);
        HIP_TRY(hipStreamWaitEvent(gen_st, E.fitted, 0));
        const double *znorm = nullptr;
        if (R.stored_normals) {
            PROF_LAUNCH(gen_st, "k_eval_normals",
                        launch_eval_normals(gen_st, t, R.ite, static_cast<uint32_t>(R.items_e.size()),
                                            static_cast<double *>(R.znorm_d),
                                            static_cast<unsigned long long *>(R.fail_e_d)));
            znorm = static_cast<const double *>(R.znorm_d);
        }
        // the escrowed instantiation also serves spot-model groups
        PROF_LAUNCH(gen_st, "k_evaluate",
                    launch_evaluate(gen_st, spec.fused, spec.greeks, R.escrow, tile, t, fp, R.ite,
                                    static_cast<uint32_t>(R.items_e.size()), R.max_strikes, R.max_obs, models,
                                    static_cast<double *>(R.part_e_d), static_cast<uint32_t *>(R.counts_d),
                                    static_cast<uint32_t *>(R.nosup_d), static_cast<RowError *>(R.errors_d),
                                    static_cast<unsigned long long *>(R.fail_e_d), znorm));
        PROF_LAUNCH(gen_st, "k_eval_reduce",
                    launch_eval_reduce(gen_st, t, R.rows, R.nrows, static_cast<double *>(R.part_e_d), greeks ? 1 : 0,
                                       static_cast<double *>(R.eval_out_d)));
        PROF_LAUNCH(gen_st, "k_row_summary",
                    launch_row_summary(gen_st, t, models, static_cast<uint32_t *>(R.counts_d),
                                       static_cast<uint32_t *>(R.nosup_d), R.nrows,
                                       static_cast<RowSummary *>(R.summary_d)));
        // results: DMA transfers into the staging area, copied out once the batch is done
        size_t off = R.tables_bytes;
        for (const auto &d : R.downloads()) {
            if (d.bytes) HIP_TRY(hipMemcpyAsync(R.staging + off, *d.dev, d.bytes, hipMemcpyDeviceToHost, gen_st));
            off += (d.bytes + 255) & ~size_t(255);
        }
        HIP_TRY(hipEventRecord(E.done, gen_st));
        return FOHIP_OK;
    };

    // ---- results of a finished batch: failures in the reference's processing order, row outputs
    auto finish = [&](BatchRun &R) -> int {
        HIP_TRY(hipEventSynchronize(ex->slot[R.io].done));
        HIP_TRY(hipGetLastError());
        {
            size_t off = R.tables_bytes;
            for (const auto &d : R.downloads()) {
                if (d.bytes) std::memcpy(d.host, R.staging + off, d.bytes);
                off += (d.bytes + 255) & ~size_t(255);
            }
        }
        if (ex->profiling) ex->collect();
        if (R.sweep && R.sweep_abort != 0u)
            return fail(err, FOHIP_E_HIP, FOHIP_PHASE_FIT,
                        "fit sweep: a team of blocks did not gather within the wait limit (%u blocks launched, "
                        "largest team %u); the results of this request are void",
                        R.sweep_slots, R.sw.most_members);
        const DevTables &t = R.t;
        FitParams fp = spec.fp;
        // Failures, in the reference's processing order: groups in order; within a group the
        // training simulation, then strike rows' fits in row order, then the evaluation
        // simulation, then strike rows' evaluations.
        for (uint32_t g = 0; g < R.ngroups; ++g) {
            const DevGroup &G = R.gt.groups[g];
            const GroupPlan &P = R.plans[g];
            const int64_t first_row = static_cast<int64_t>(R.row0 + P.strike_begin);
            auto sim_failure = [&](int evaluation, unsigned long long path) -> int {
                // the replay runs alone on the fit stream once every batch in flight has drained
                HIP_TRY(hipStreamSynchronize(gen_st));
                HIP_TRY(hipStreamSynchronize(fit_st));
                std::vector<double> info(4, 0.0);
                HIP_TRY(launch_replay_path(fit_st, t, g, evaluation, static_cast<uint32_t>(path),
                                           static_cast<double *>(R.replay_d)));
                HIP_TRY(hipMemcpyAsync(info.data(), R.replay_d, 4 * sizeof(double), hipMemcpyDeviceToHost, fit_st));
                HIP_TRY(hipStreamSynchronize(fit_st));
                const int phase = evaluation ? FOHIP_PHASE_EVALUATION_SIMULATION : FOHIP_PHASE_TRAINING_SIMULATION;
                const char *role = evaluation ? "evaluation" : "training";
                const uint32_t interval = static_cast<uint32_t>(info[0] < 0 ? 0 : info[0]);
                set_fields(err, first_row, interval + 1, static_cast<int64_t>(path), interval, info[1], info[2]);
                if (info[3] != 0.0)
                    return fail(err, FOHIP_E_NORMAL_REJECTION_LIMIT, phase,
                                "LSM %s simulation failed: random-normal rejection limit reached at interval %u, "
                                "path %llu; no variate or fallback was produced",
                                role, interval, path);
                return fail(err, FOHIP_E_NON_FINITE_TRANSITION, phase,
                            "LSM %s simulation failed: GBM path transition from node %u to %u at path %llu produced "
                            "non-finite spot from state %.17g and variate %.17g",
                            role, interval, interval + 1, path, info[1], info[2]);
            };
            if (R.fail_t[g] != kNoError) return sim_failure(0, R.fail_t[g]);
            if (R.bad_group == static_cast<int64_t>(g)) {
                set_fields(err, first_row, R.bad_node, -1, -1, 0, 0);
                return fail(err, FOHIP_E_UNREPRESENTABLE_DISCOUNT, FOHIP_PHASE_FIT,
                            "discount factor to observation node %lld is not finite and positive",
                            (long long)R.bad_node);
            }
            for (uint32_t s = P.strike_begin; s < P.strike_end; ++s) {
                const RowError &e = R.row_err[s];
                if (e.key == kNoError) continue;
                const uint32_t step = static_cast<uint32_t>(e.key >> 44);
                if (step >= kEvalStepBase) continue;
                const uint32_t stage = static_cast<uint32_t>((e.key >> 36) & 0xFF);
                const uint64_t path = e.key & ((1ull << 36) - 1);
                const int64_t row = static_cast<int64_t>(R.row0 + s);
                uint32_t node = 0;
                if (step < G.fit_rows) node = P.obs_nodes[G.obs - 2 - step];
                if (stage == kStageRegression) {
                    char msg[256];
                    set_fields(err, row, node, -1, e.index, e.observed, e.limit);
                    regression_message(e.code, msg, sizeof msg, fp.degree, e);
                    return fail(err, e.code, FOHIP_PHASE_FIT, "continuation regression failed at exercise node %u: %s",
                                node, msg);
                }
                set_fields(err, row, node, static_cast<int64_t>(path), -1, 0, 0);
                if (stage == kStageApply)
                    return fail(err, FOHIP_E_NON_FINITE_PREDICTION, FOHIP_PHASE_FIT,
                                "continuation prediction failed at exercise node %u, path %llu", node,
                                (unsigned long long)path);
                if (stage == kStagePredictor)
                    return fail(err, FOHIP_E_NON_FINITE_PREDICTOR, FOHIP_PHASE_FIT,
                                "continuation regression failed at exercise node %u: regression predictor must be "
                                "finite at path %llu",
                                node, (unsigned long long)path);
                return fail(err, FOHIP_E_NON_FINITE_DISCOUNTED_PAYOFF, FOHIP_PHASE_FIT,
                            "LSM evaluation produced non-finite discounted payoff at path %llu", (unsigned long long)path);
            }
            if (R.fail_e[g] != kNoError) return sim_failure(1, R.fail_e[g]);
            for (uint32_t s = P.strike_begin; s < P.strike_end; ++s) {
                const RowError &e = R.row_err[s];
                const RowSummary &rs = R.summary[s];
                const int64_t row = static_cast<int64_t>(R.row0 + s);
                int64_t err_row_obs = -1;
                uint32_t stage = 0;
                uint64_t path = 0;
                if (e.key != kNoError && (e.key >> 44) >= kEvalStepBase) {
                    err_row_obs = static_cast<int64_t>((e.key >> 44) - kEvalStepBase);
                    stage = static_cast<uint32_t>((e.key >> 36) & 0xFF);
                    path = e.key & ((1ull << 36) - 1);
                }
                if (rs.nosupport_row >= 0 && (err_row_obs < 0 || rs.nosupport_row <= err_row_obs)) {
                    const uint32_t node = P.obs_nodes[rs.nosupport_row];
                    set_fields(err, row, node, -1, rs.nosupport_count, 0, 0);
                    return fail(err, FOHIP_E_OUTSIDE_TRAINING_SUPPORT, FOHIP_PHASE_EVALUATION,
                                "fixed policy has no training support at exercise node %u, but %u active evaluation "
                                "paths are in the money",
                                node, rs.nosupport_count);
                }
                if (err_row_obs >= 0) {
                    const uint32_t node = P.obs_nodes[err_row_obs];
                    set_fields(err, row, node, static_cast<int64_t>(path), -1, 0, 0);
                    if (stage == kStageEvalPrediction)
                        return fail(err, FOHIP_E_NON_FINITE_PREDICTION, FOHIP_PHASE_EVALUATION,
                                    "continuation prediction failed at exercise node %u, path %llu", node,
                                    (unsigned long long)path);
                    return fail(err, FOHIP_E_NON_FINITE_DISCOUNTED_PAYOFF, FOHIP_PHASE_EVALUATION,
                                "LSM evaluation produced non-finite discounted payoff at path %llu",
                                (unsigned long long)path);
                }
            }
        }
        // ---- row outputs
        for (uint32_t s = 0; s < R.nrows; ++s) {
            const uint32_t g = R.strike_group[s];
            const GroupPlan &P = R.plans[g];
            RowOut &o = out[R.row0 + s];
            o.sum = R.summary[s];
            o.ties_training = R.ties_h[s] + static_cast<uint64_t>(R.summary[s].ties_training);
            o.ties_evaluation = static_cast<uint64_t>(R.eval_out[static_cast<size_t>(s) * 16 + 15]);
            const double n = static_cast<double>(P.ne);
            if (R.summary[s].origin_fitted && R.summary[s].origin_exercised) {
                // policy.rs:672-682: deterministic origin exercise
                o.price = std::max(P.S0 - R.bstrikes[s], 0.0);
                o.sd = o.se = 0.0;
                o.has_sd = true;
                o.sum.early = static_cast<uint32_t>(P.ne);
                o.sum.expiry = 0;
                o.greek[0] = P.S0 > R.bstrikes[s] ? 1.0 : 0.0;
                continue;
            }
            const double *e = &R.eval_out[static_cast<size_t>(s) * 16];
            o.price = e[0];
            if (P.ne > 1) {
                o.sd = std::sqrt(e[1] / (n - 1.0));
                o.se = o.sd / std::sqrt(n);
                o.has_sd = true;
            }
            for (int a = 0; a < 4; ++a) {
                o.greek[a] = e[3 + a];
                o.greek_se[a] = e[7 + a];
            }
        }
        // append with strike rows in batch order: models are [strike][obs] per group
        if (models_out) models_out->insert(models_out->end(), R.models_h.begin(), R.models_h.end());
        if (counts_out) counts_out->insert(counts_out->end(), R.counts_h.begin(), R.counts_h.end());
        // keep the last batch's layout for diagnostics
        ex->last_models = R.models_h;
        return FOHIP_OK;
    };

    int rc = FOHIP_OK;
    if (!overlapped) {
        // The host builds batch b + 1 and queues it behind batch b before it waits for b's
        // results, so the device goes from one batch to the next without waiting for the host.
        // Consecutive batches use the same device buffers, ordered by the generation stream (b + 1's
        // upload and simulation follow b's evaluation, which follows b's backward induction), and
        // alternate between the two slots' events and page-locked staging. The fixed region of the
        // allocation-free entry point has one staging area: there b + 1 is queued after b is done.
        const bool pipelined = !fixed;
        std::unique_ptr<BatchRun> cur = prepare(0);
        if ((rc = issue_front(*cur)) || (rc = issue_back(*cur))) return rc;
        for (size_t b = 0; b < batches.size(); ++b) {
            std::unique_ptr<BatchRun> next = b + 1 < batches.size() ? prepare(b + 1) : nullptr;
            int rc_next = FOHIP_OK;
            if (next && pipelined && !(rc_next = issue_front(*next))) rc_next = issue_back(*next);
            // a failure to queue b + 1 is reported after batch b's own results, which precede it
            if ((rc = finish(*cur))) return rc;
            if (rc_next) return rc_next;
            if (next && !pipelined && ((rc = issue_front(*next)) || (rc = issue_back(*next)))) return rc;
            cur = std::move(next);
        }
    } else {
        std::unique_ptr<BatchRun> prev, cur = prepare(0);
        if ((rc = issue_front(*cur))) return rc;
        for (size_t b = 0; b < batches.size(); ++b) {
            std::unique_ptr<BatchRun> next;
            int rc_next = FOHIP_OK;
            if (b + 1 < batches.size()) {
                next = prepare(b + 1);   // host work while the device runs batches b - 1 and b
                if (prev) {              // b - 1 holds the slot b + 1 is about to use
                    if ((rc = finish(*prev))) return rc;
                    prev.reset();
                }
                rc_next = issue_front(*next);
            }
            // a failure to queue b + 1 is reported after batch b's own results, which precede it
            if ((rc = issue_back(*cur))) return rc;
            if (rc_next) {
                if (prev && (rc = finish(*prev))) return rc;
                if ((rc = finish(*cur))) return rc;
                return rc_next;
            }
            if (prev && (rc = finish(*prev))) return rc;
            prev = std::move(cur);
            cur = std::move(next);
        }
        if (prev && (rc = finish(*prev))) return rc;
    }
    if (ex->profiling) {
        HIP_TRY(hipStreamSynchronize(fit_st));
        ex->collect();
    }
    return FOHIP_OK;
}

// ============================================================================ single call

static int parse_regression(uint32_t degree, double rank_tol, double cond_limit, int32_t tie, int32_t support,
                            FitParams &fp, fohip_error *err) {
    // regression.rs:23-90
    if (degree > 4)
        return fail(err, FOHIP_E_UNSUPPORTED_DEGREE, 0, "polynomial degree %u exceeds the supported maximum 4", degree);
    if (!std::isfinite(rank_tol) || rank_tol <= 0.0 || rank_tol >= 1.0)
        return fail(err, FOHIP_E_INVALID_REGRESSION_SPEC, 0,
                    "rank tolerance must be finite and strictly between 0 and 1; observed %.17g", rank_tol);
    const double maximum = 1.0 / rank_tol;
    if (!std::isfinite(cond_limit) || cond_limit < 1.0 || cond_limit > maximum)
        return fail(err, FOHIP_E_INVALID_REGRESSION_SPEC, 0,
                    "condition limit must be finite and in [1, %.17g]; observed %.17g", maximum, cond_limit);
    if (tie != 0 && tie != 1) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown exercise tie rule %d", tie);
    if (support != 0 && support != 1)
        return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown insufficient-support policy %d", support);
    fp.degree = static_cast<int32_t>(degree);
    fp.tie = tie;
    fp.support = support;
    fp.pad = 0;
    fp.rank_tol = rank_tol;
    fp.cond_limit = cond_limit;
    return FOHIP_OK;
}

static bool same_grid(const double *a, size_t na, const double *b, size_t nb) {
    if (na != nb) return false;
    for (size_t i = 0; i < na; ++i) {
        uint64_t x, y;
        std::memcpy(&x, &a[i], 8);
        std::memcpy(&y, &b[i], 8);
        if (x != y) return false;
    }
    return true;
}

// fohip_price_lsm_call's body; `fixed` (the allocation-free entry point) carves the device
// buffers from a fixed region instead of the executor's own.
static int price_call(fohip_executor *ex, const fohip_call_request *rq, fohip_call_result *res, fohip_error *err,
                      Workspace *fixed) {
    int rc;
    // lib.rs:1170-1248 parse order: financial model, regression spec, estimator
    if ((rc = fohip_validate_event_schedule(rq->schedule_times, rq->schedule_nodes, rq->dividend_nodes,
                                            rq->dividend_amounts, rq->dividends, rq->exercise_nodes,
                                            rq->exercise_count, rq->expiry_node, err)))
        return rc;
    if ((rc = fohip_validate_log_discount_curve(rq->curve_times, rq->log_discounts, rq->curve_nodes, err))) return rc;
    if ((rc = fohip_validate_cumulative_yield_curve(rq->yield_times, rq->cumulative_yields, rq->yield_nodes, err)))
        return rc;
    if ((rc = fohip_validate_constant_volatility(rq->volatility, err))) return rc;
    if (rq->dividend_policy != 0 && rq->dividend_policy != 1)
        return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown dividend policy %d", rq->dividend_policy);
    if (!same_grid(rq->schedule_times, rq->schedule_nodes, rq->curve_times, rq->curve_nodes))
        return fail(err, FOHIP_E_TIME_GRID_MISMATCH, 0, "discount curve and event schedule must use the identical time grid");
    if (!same_grid(rq->schedule_times, rq->schedule_nodes, rq->yield_times, rq->yield_nodes))
        return fail(err, FOHIP_E_TIME_GRID_MISMATCH, 0,
                    "continuous dividend-yield curve and event schedule must use the identical time grid");
    FitParams fp;
    if ((rc = parse_regression(rq->polynomial_degree, rq->rank_tolerance, rq->condition_limit, rq->tie_rule,
                               rq->insufficient_support, fp, err)))
        return rc;
    if (!std::isfinite(rq->strike) || rq->strike <= 0.0)
        return fail(err, FOHIP_E_INVALID_STRIKE, 0, "call strike must be finite and strictly positive; observed %.17g",
                    rq->strike);
    // policy.rs:306-336
    if (rq->training_domain == rq->evaluation_domain)
        return fail(err, FOHIP_E_RANDOM_DOMAIN_COLLISION, 0,
                    "LSM training and evaluation random domains must differ; both use domain %u", rq->training_domain);
    if (!canonical_zero_or_positive(rq->initial_spot))
        return fail(err, FOHIP_E_INVALID_INITIAL_SPOT, FOHIP_PHASE_TRAINING_SIMULATION,
                    "LSM training simulation failed: initial spot must be finite, nonnegative, and canonical; observed "
                    "%.17g",
                    rq->initial_spot);
    if (rq->training_paths == 0 || rq->evaluation_paths == 0)
        return fail(err, FOHIP_E_EMPTY_PATH_SET, rq->training_paths == 0 ? 1 : 3,
                    "LSM %s simulation failed: simulation path count must be positive",
                    rq->training_paths == 0 ? "training" : "evaluation");
    if (rq->training_paths > 0xFFFFFFFFull || rq->evaluation_paths > 0xFFFFFFFFull)
        return fail(err, FOHIP_E_PATH_NAMESPACE_OVERFLOW, 0, "path count exceeds the 32-bit path namespace");
    if (rq->training_paths >= (1ull << 36) / 16 || rq->evaluation_paths >= (1ull << 36) / 16)
        return fail(err, FOHIP_E_UNSUPPORTED, 0, "path count exceeds this engine's per-group limit");
    GroupPlan p;
    p.times.assign(rq->schedule_times, rq->schedule_times + rq->schedule_nodes);
    p.logd.assign(rq->log_discounts, rq->log_discounts + rq->curve_nodes);
    p.q.assign(rq->cumulative_yields, rq->cumulative_yields + rq->yield_nodes);
    p.cash.assign(p.times.size(), 0.0);
    p.has_cash.assign(p.times.size(), 0);
    for (size_t i = 0; i < rq->dividends; ++i) {
        p.cash[rq->dividend_nodes[i]] = rq->dividend_amounts[i];
        p.has_cash[rq->dividend_nodes[i]] = 1;
    }
    p.obs_nodes.assign(rq->exercise_nodes, rq->exercise_nodes + rq->exercise_count);
    p.obs_nodes.push_back(rq->expiry_node);
    p.S0 = rq->initial_spot;
    p.sigma = rq->volatility;
    p.policy = rq->dividend_policy;
    p.strike_begin = 0;
    p.strike_end = 1;
    p.nt = rq->training_paths;
    p.ne = rq->evaluation_paths;
    p.train_seed = rq->training_seed;
    p.eval_seed = rq->evaluation_seed;
    p.train_domain = rq->training_domain;
    p.eval_domain = rq->evaluation_domain;
    RunSpec spec;
    spec.fp = fp;
    spec.greeks = rq->greeks;
    set_method(spec, ex->precision);
    std::vector<RowOut> out;
    std::vector<Model> models;
    std::vector<uint32_t> counts;
    std::vector<GroupPlan> plans{p};
    std::vector<double> strikes{rq->strike};
    rc = run_groups(ex, plans, strikes, spec, out, &models, &counts, err, fixed);
    if (rc) return rc;
    const RowOut &o = out[0];
    const uint32_t obs = static_cast<uint32_t>(p.obs_nodes.size());
    res->price = o.price;
    res->paths = static_cast<uint32_t>(p.ne);
    res->has_standard_error = o.has_sd ? 1 : 0;
    res->sample_standard_deviation = o.sd;
    res->standard_error = o.se;
    res->observation_rows = obs;
    const bool origin_exercise = o.sum.origin_fitted && o.sum.origin_exercised;
    for (uint32_t r = 0; r < obs; ++r) {
        if (res->observation_nodes) res->observation_nodes[r] = p.obs_nodes[r];
        if (res->exercise_counts) res->exercise_counts[r] = origin_exercise ? (r == 0 ? res->paths : 0) : counts[r];
    }
    for (int a = 0; a < 4; ++a) {   // unrequested sensitivities are not computed
        const bool wanted = (rq->greeks & (1u << a)) != 0;
        res->greeks[a] = wanted ? o.greek[a] : std::nan("");
        res->greek_standard_errors[a] = wanted ? o.greek_se[a] : std::nan("");
    }
    res->near_tie_training_decisions = o.ties_training;
    res->near_tie_evaluation_decisions = o.ties_evaluation;
    if (res->diagnostics) {
        for (uint32_t r = 0; r + 1 < obs; ++r) {
            const Model &m = models[r];
            fohip_node_diagnostic &d = res->diagnostics[r];
            std::memset(&d, 0, sizeof d);
            d.node = p.obs_nodes[r];
            switch (m.kind) {
                case kRuleOrigin:
                    d.kind = 0;
                    d.exercise = m.exercise;
                    d.estimated_continuation = m.estimate;
                    break;
                case kRuleNoSupport:
                    d.kind = 1;
                    break;
                case kRuleContinue:
                    d.kind = 2;
                    d.in_the_money_paths = m.itm;
                    d.columns = static_cast<uint32_t>(fp.degree + 1);
                    break;
                default:
                    d.kind = 3;
                    d.in_the_money_paths = m.itm;
                    d.observations = m.itm;
                    d.columns = static_cast<uint32_t>(fp.degree + 1);
                    d.rank = static_cast<uint32_t>(m.rank);
                    d.condition_number = m.condition;
                    d.predictor_mean = m.mean;
                    d.predictor_scale = m.scale;
                    for (int k = 0; k < 4; ++k) {
                        d.alpha[k] = m.alpha[k];
                        d.beta[k] = m.beta[k];
                    }
                    for (int k = 0; k < 5; ++k) d.coefficients[k] = m.coef[k];
            }
        }
    }
    return FOHIP_OK;
}

extern "C" int fohip_price_lsm_call(fohip_executor *ex, const fohip_call_request *rq, fohip_call_result *res,
                                    fohip_error *err) {
    clear_error(err);
    if (!ex || !rq || !res) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null executor, request or result");
    std::lock_guard<std::mutex> lock(ex->mu);
    return price_call(ex, rq, res, err, nullptr);
}

// One lock per device: every allocation-free call on a device uses that device's one region.
static std::mutex g_noalloc_mu[kNoAllocMaxDevices];

extern "C" int fohip_price_lsm_call_noalloc(fohip_executor *ex, const fohip_call_request *rq, fohip_call_result *res,
                                            fohip_error *err) {
    clear_error(err);
    if (!ex || !rq || !res) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null executor, request or result");
    if (ex->device < 0 || ex->device >= kNoAllocMaxDevices)
        return fail(err, FOHIP_E_UNSUPPORTED, 0, "the allocation-free entry point supports devices 0..%d",
                    kNoAllocMaxDevices - 1);
    if (ex->profiling)
        return fail(err, FOHIP_E_UNSUPPORTED, 0, "the allocation-free entry point needs per-kernel profiling off");
    std::lock_guard<std::mutex> device_lock(g_noalloc_mu[ex->device]);
    std::lock_guard<std::mutex> lock(ex->mu);
    HIP_TRY(hipSetDevice(ex->device));
    void *region = nullptr;
    size_t bytes = 0;
    HIP_TRY(noalloc_scratch(ex->device, &region, &bytes));
    Workspace fixed(ex, region, bytes);
    return price_call(ex, rq, res, err, &fixed);
}

// =========================================================================== grouped batch

namespace {

int total_cmp(double a, double b) { return a < b ? -1 : (a > b ? 1 : 0); }

int64_t search_time(const std::vector<double> &times, double t) {
    auto it = std::lower_bound(times.begin(), times.end(), t, [](double a, double b) { return total_cmp(a, b) < 0; });
    if (it != times.end() && total_cmp(*it, t) == 0) return it - times.begin();
    return -1;
}

}  // namespace

// validate_escrowed_row (input.rs:798-832) with escrowed_origin_prepaid_cash (input.rs:834-860):
// F0 = S0 - sum_k D_k exp(-r t_k) (schedule order) must be finite and positive, and a spot-matched
// volatility sigma S0 / F0 finite and positive. `sigma_f` receives the path volatility
// (path_volatility, pricing.rs:840-860).
static int validate_escrowed_group(size_t g, double S0, double rate, double sigma, const double *times,
                                   const double *amounts, size_t count, int32_t convention, double *sigma_f,
                                   fohip_error *err) {
    double prepaid_cash = 0.0;
    for (size_t e = 0; e < count; ++e) {
        const double discounted = amounts[e] * std::exp(-rate * times[e]);
        prepaid_cash += discounted;
        if (!std::isfinite(discounted) || !std::isfinite(prepaid_cash))
            return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0,
                        "escrowed prepaid cash is not finite at batch row %zu, event %zu: time=%.17g, amount=%.17g, "
                        "rate=%.17g",
                        g, e, times[e], amounts[e], rate);
    }
    const double prepaid_forward = S0 - prepaid_cash;
    if (!std::isfinite(prepaid_forward) || prepaid_forward <= 0.0)
        return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0,
                    "escrowed prepaid forward must be strictly positive at batch row %zu; initial spot is %.17g and "
                    "origin prepaid cash is %.17g",
                    g, S0, prepaid_cash);
    double path_sigma = sigma;
    if (convention == FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN) {
        path_sigma = sigma * S0 / prepaid_forward;
        if (!std::isfinite(path_sigma) || path_sigma <= 0.0)
            return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0,
                        "escrowed spot-matched volatility is invalid at batch row %zu: supplied=%.17g, initial_spot=%.17g, "
                        "prepaid_forward=%.17g, transformed=%.17g",
                        g, sigma, S0, prepaid_forward, path_sigma);
    }
    if (sigma_f) *sigma_f = path_sigma;
    return FOHIP_OK;
}

// EscrowedGbmPathPlan::try_new (path.rs:628-666) with escrowed_prepaid_cash (path.rs:897-944):
// prepaid[node] = prepaid[node + 1] exp(logD[node + 1] - logD[node]) + same-node cash (expiry
// holds a terminal amount), prepaid[0] replaced by sum_k D_k exp(logD[node_k]) in schedule order.
// Also the rate derivative of every prepaid value and the Greek chain factors of the observed
// stock through F0 = S0 - prepaid[0] and, spot matched, sigma_F = sigma S0 / F0.
static int escrowed_plan(GroupPlan &p, const std::vector<double> &times, const std::vector<uint32_t> &cash_node,
                         const std::vector<double> &cash_a, double S0, double sigma, int32_t convention,
                         size_t group_row, fohip_error *err) {
    const size_t nodes = times.size();
    p.escrowed = true;
    p.prepaid.assign(nodes, 0.0);
    p.prepaid_dr.assign(nodes, 0.0);
    size_t event = cash_a.size();
    const size_t expiry = nodes - 1;
    if (event > 0 && cash_node[event - 1] == expiry) {
        event -= 1;
        p.prepaid[expiry] = cash_a[event];
    }
    for (size_t k = nodes - 1; k-- > 0;) {
        const double interval_log_discount = p.logd[k + 1] - p.logd[k];
        const double factor = std::exp(interval_log_discount);
        const double discounted_future = p.prepaid[k + 1] * factor;
        double same_node_cash = 0.0;
        if (event > 0 && cash_node[event - 1] == k) {
            event -= 1;
            same_node_cash = cash_a[event];
        }
        const double value = discounted_future + same_node_cash;
        if (!std::isfinite(value))
            return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0,
                        "row %zu escrowed prepaid cash is not finite at node %zu", group_row, k);
        p.prepaid[k] = value;
        // d factor / dr = -(t[k + 1] - t[k]) factor (logD = -r t)
        p.prepaid_dr[k] = p.prepaid_dr[k + 1] * factor - p.prepaid[k + 1] * factor * (times[k + 1] - times[k]);
    }
    double origin = 0.0, origin_dr = 0.0;
    for (size_t e = 0; e < cash_a.size(); ++e) {
        const double pv = cash_a[e] * std::exp(p.logd[cash_node[e]]);
        origin += pv;
        origin_dr += -times[cash_node[e]] * pv;
        if (!std::isfinite(origin))
            return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0, "row %zu escrowed prepaid cash is not finite at node 0",
                        group_row);
    }
    p.prepaid[0] = origin;
    p.prepaid_dr[0] = origin_dr;
    p.prepaid_dr[expiry] = 0.0;
    const double F0 = S0 - origin;
    if (!std::isfinite(F0) || F0 <= 0.0)
        return fail(err, FOHIP_E_ESCROWED_PREPAID_FORWARD, 0,
                    "escrowed prepaid forward must be strictly positive at batch row %zu; initial spot is %.17g and "
                    "origin prepaid cash is %.17g",
                    group_row, S0, origin);
    p.start = F0;
    p.sigma = convention == FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN ? sigma * S0 / F0 : sigma;
    p.chain[0] = convention == FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN ? -sigma * origin / (F0 * F0) : 0.0;
    p.chain[1] = convention == FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN ? S0 / F0 : 1.0;
    p.chain[2] = -origin_dr;
    p.chain[3] = convention == FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN ? sigma * S0 * origin_dr / (F0 * F0) : 0.0;
    return FOHIP_OK;
}

static int build_grouped_plans(const fohip_grouped_request *rq, std::vector<GroupPlan> &plans,
                               std::vector<double> &strikes, fohip_error *err) {
    const size_t groups = rq->groups;
    // input.rs:617-723
    if (groups == 0 || rq->rows == 0) return fail(err, FOHIP_E_EMPTY_BATCH, 0, "batch must contain at least one contract");
    if (!rq->group_offsets || !rq->initial_spots || !rq->maturities || !rq->rates || !rq->volatilities || !rq->strikes)
        return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null batch column");
    // validate_group_offsets (input.rs:761-796)
    if (rq->group_offsets[0] != 0)
        return fail(err, FOHIP_E_GROUP_OFFSETS, 0, "group offsets must start at zero; observed %u", rq->group_offsets[0]);
    for (size_t g = 0; g < groups; ++g)
        if (rq->group_offsets[g] >= rq->group_offsets[g + 1])
            return fail(err, FOHIP_E_GROUP_OFFSETS, 0, "group %zu has an empty or decreasing range %u..%u", g,
                        rq->group_offsets[g], rq->group_offsets[g + 1]);
    if (rq->group_offsets[groups] != rq->rows)
        return fail(err, FOHIP_E_GROUP_OFFSETS, 0, "final group offset %u must equal the strike row count %zu",
                    rq->group_offsets[groups], rq->rows);
    for (size_t r = 0; r < rq->rows; ++r)
        if (!std::isfinite(rq->strikes[r]) || rq->strikes[r] <= 0.0)
            return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0, "strikes[%zu] must be finite and strictly positive; observed %.17g",
                        r, rq->strikes[r]);
    if (rq->dividend_model < FOHIP_DIVIDENDS_CONTINUOUS_YIELD || rq->dividend_model > FOHIP_DIVIDENDS_ESCROWED_CASH)
        return fail(err, FOHIP_E_UNSUPPORTED, 0, "unsupported dividend model %d", rq->dividend_model);
    const bool has_cash_schedule = rq->dividend_model != FOHIP_DIVIDENDS_CONTINUOUS_YIELD;
    const bool escrowed = rq->dividend_model == FOHIP_DIVIDENDS_ESCROWED_CASH;
    if (has_cash_schedule) {
        if (!rq->cash_offsets) return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "cash-dividend offsets are missing");
        int rc = fohip_validate_ragged_offsets(rq->cash_offsets, groups + 1, rq->cash_events, err);
        if (rc) return rc;
        if (rq->maturity_boundary != 0 && rq->maturity_boundary != 1)
            return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "unsupported cash maturity boundary");
        if (!escrowed && rq->cash_policy != 0 && rq->cash_policy != 1)
            return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown dividend policy %d", rq->cash_policy);
        if (escrowed && rq->escrowed_volatility != FOHIP_ESCROWED_PREPAID_FORWARD &&
            rq->escrowed_volatility != FOHIP_ESCROWED_SPOT_MATCHED_AT_ORIGIN)
            return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown escrowed volatility convention %d",
                        rq->escrowed_volatility);
    }
    for (size_t g = 0; g < groups; ++g) {
        const double S0 = rq->initial_spots[g], T = rq->maturities[g], r = rq->rates[g], v = rq->volatilities[g];
        if (!canonical_zero_or_positive(S0))
            return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0, "initial_spots[%zu] must be finite and nonnegative; observed %.17g",
                        g, S0);
        if (!std::isfinite(T) || T <= 0.0)
            return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0, "maturities[%zu] must be finite and strictly positive; observed %.17g",
                        g, T);
        if (!std::isfinite(r))
            return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0, "rates[%zu] must be finite; observed %.17g", g, r);
        if (!std::isfinite(v) || v <= 0.0)
            return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0,
                        "volatilities[%zu] must be finite and strictly positive; observed %.17g", g, v);
        if (rq->dividend_model == 0) {
            const double q = rq->dividend_yields[g];
            if (!canonical_zero_or_positive(q))
                return fail(err, FOHIP_E_INVALID_ROW_VALUE, 0,
                            "continuous_dividend_yields[%zu] must be finite and nonnegative; observed %.17g", g, q);
        } else {
            // validate_cash_schedule (input.rs:725-759)
            const int64_t cb = rq->cash_offsets[0];
            for (int64_t e = rq->cash_offsets[g] - cb; e < rq->cash_offsets[g + 1] - cb; ++e) {
                const double t = rq->cash_times[e];
                if (!std::isfinite(t) || t <= 0.0)
                    return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "cash dividend time must be finite and positive");
                if (t > T)
                    return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "row %zu cash event %lld at %.17g is after maturity %.17g",
                                g, (long long)(e - (rq->cash_offsets[g] - cb)), t, T);
                if (t == T && rq->maturity_boundary == FOHIP_BOUNDARY_STRICTLY_BEFORE)
                    return fail(err, FOHIP_E_CASH_SCHEDULE, 0,
                                "row %zu cash event %lld lies at excluded maturity %.17g", g,
                                (long long)(e - (rq->cash_offsets[g] - cb)), T);
                if (e > rq->cash_offsets[g] - cb && rq->cash_times[e - 1] >= t)
                    return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "row %zu cash event times must be strictly increasing", g);
                int rc = fohip_validate_cash_dividend_amount(rq->cash_amounts[e], err);
                if (rc) return rc;
            }
            if (escrowed) {
                const int64_t first = rq->cash_offsets[g] - cb, count = rq->cash_offsets[g + 1] - cb - first;
                int rc = validate_escrowed_group(g, S0, r, v, rq->cash_times + first, rq->cash_amounts + first,
                                                 static_cast<size_t>(count), rq->escrowed_volatility, nullptr, err);
                if (rc) return rc;
            }
        }
    }
    // method-level checks (lib.rs:1302-1400, input.rs:201-238)
    auto nt_of = [&](size_t g) -> uint64_t {
        return rq->training_path_counts ? rq->training_path_counts[g] : rq->training_paths;
    };
    auto ne_of = [&](size_t g) -> uint64_t {
        return rq->evaluation_path_counts ? rq->evaluation_path_counts[g] : rq->evaluation_paths;
    };
    for (size_t g = 0; g < groups; ++g) {
        if (nt_of(g) == 0 || nt_of(g) > 0xFFFFFFFFull)
            return fail(err, FOHIP_E_INVALID_PATH_COUNT, 0, "training path count of group %zu must be in [1, 4294967295]",
                        g);
        if (ne_of(g) < 2 || ne_of(g) > 0xFFFFFFFFull)
            return fail(err, FOHIP_E_INSUFFICIENT_EVALUATION_PATHS, 0,
                        "batched LSM requires at least two independent evaluation paths (group %zu)", g);
        if (nt_of(g) >= (1ull << 36) / 16 || ne_of(g) >= (1ull << 36) / 16)
            return fail(err, FOHIP_E_UNSUPPORTED, 0, "path count of group %zu exceeds this engine's per-group limit", g);
    }
    {
        const uint64_t te = static_cast<uint64_t>(rq->training_domain_start) + groups;
        const uint64_t ee = static_cast<uint64_t>(rq->evaluation_domain_start) + groups;
        if (te > (1ull << 32) || ee > (1ull << 32))
            return fail(err, FOHIP_E_RANDOM_DOMAIN_RANGE, 0, "random domain range overflows the 32-bit domain lane");
        if (rq->training_domain_start < ee && rq->evaluation_domain_start < te)
            return fail(err, FOHIP_E_RANDOM_DOMAIN_RANGE, 0,
                        "training domains %u..%llu and evaluation domains %u..%llu overlap", rq->training_domain_start,
                        (unsigned long long)te, rq->evaluation_domain_start, (unsigned long long)ee);
    }
    if (rq->mesh_kind == FOHIP_MESH_UNIFORM && rq->base_intervals == 0)
        return fail(err, FOHIP_E_EXERCISE_MESH, 0, "uniform exercise mesh needs at least one base interval");
    if (rq->mesh_kind == FOHIP_MESH_CASH_DIVIDEND_DATES && !has_cash_schedule)
        return fail(err, FOHIP_E_EXERCISE_MESH, 0,
                    "cash-dividend-date exercise mesh requires a deterministic cash-dividend model");
    if (rq->mesh_kind == FOHIP_MESH_EXPLICIT) {
        if (!rq->exercise_offsets) return fail(err, FOHIP_E_EXERCISE_MESH, 0, "explicit exercise offsets are missing");
        int rc = fohip_validate_ragged_offsets(rq->exercise_offsets, groups + 1, rq->exercise_events, err);
        if (rc) return rc;
    }
    if (rq->mesh_kind < 0 || rq->mesh_kind > 2) return fail(err, FOHIP_E_EXERCISE_MESH, 0, "unknown exercise mesh");
    strikes.assign(rq->strikes, rq->strikes + rq->rows);
    plans.resize(groups);
    // One group's grid, curves and schedule (batch/pricing.rs:840-965). Groups are independent, so
    // large requests build them on several host threads; the error reported is the lowest failing
    // group's, rebuilt serially so that its message is the one a serial pass would give.
    auto build_one = [&](size_t g, fohip_error *err) -> int {
        GroupPlan &p = plans[g];
        const double T = rq->maturities[g], rate = rq->rates[g];
        std::vector<double> cash_t, cash_a;
        if (has_cash_schedule) {
            const int64_t cb = rq->cash_offsets[0];
            cash_t.assign(rq->cash_times + (rq->cash_offsets[g] - cb), rq->cash_times + (rq->cash_offsets[g + 1] - cb));
            cash_a.assign(rq->cash_amounts + (rq->cash_offsets[g] - cb), rq->cash_amounts + (rq->cash_offsets[g + 1] - cb));
        }
        std::vector<double> ex_t;
        if (rq->mesh_kind == FOHIP_MESH_EXPLICIT)
            ex_t.assign(rq->exercise_times + (rq->exercise_offsets[g] - rq->exercise_offsets[0]),
                        rq->exercise_times + (rq->exercise_offsets[g + 1] - rq->exercise_offsets[0]));
        if (rq->mesh_kind == FOHIP_MESH_CASH_DIVIDEND_DATES) {
            // validate_cash_dividend_exercise_mesh_row (batch/pricing.rs:816-838)
            if (rate < 0.0) {
                set_fields(err, rq->group_offsets[g], -1, -1, -1, rate, 0.0);
                return fail(err, FOHIP_E_EXERCISE_MESH, 0,
                            "cash-dividend-date exercise mesh requires a nonnegative continuously compounded rate "
                            "at batch row %zu; observed %s",
                            (size_t)rq->group_offsets[g], rust_f64(rate).s);
            }
            for (size_t e = 0; e < cash_t.size(); ++e)
                if (cash_t[e] == T) {
                    set_fields(err, rq->group_offsets[g], -1, -1, static_cast<int64_t>(e), T, 0.0);
                    return fail(err, FOHIP_E_EXERCISE_MESH, 0,
                                "cash-dividend-date exercise mesh does not admit terminal cash at batch row %zu, "
                                "event %zu; event and maturity are %s",
                                (size_t)rq->group_offsets[g], e, rust_f64(T).s);
                }
        }
        if (rq->mesh_kind == FOHIP_MESH_EXPLICIT) {
            for (size_t e = 0; e < ex_t.size(); ++e) {
                if (!std::isfinite(ex_t[e]) || ex_t[e] <= 0.0 || (e > 0 && ex_t[e] <= ex_t[e - 1]))
                    return fail(err, FOHIP_E_EXERCISE_MESH, 0, "explicit exercise times must be positive and strictly increasing");
                if (ex_t[e] >= T)
                    return fail(err, FOHIP_E_EXERCISE_MESH, 0,
                                "explicit exercise time at batch row %zu, event %zu, time %s must be strictly before "
                                "maturity %s",
                                (size_t)rq->group_offsets[g], e, rust_f64(ex_t[e]).s, rust_f64(T).s);
            }
        }
        // exact_union_times / exact_explicit_union_times (batch/pricing.rs:967-1047, 1049-1114)
        std::vector<double> &times = p.times;
        times.clear();
        if (rq->mesh_kind == FOHIP_MESH_EXPLICIT) {
            times.push_back(0.0);
            size_t e = 0, c = 0;
            while (e < ex_t.size() || c < cash_t.size()) {
                double next;
                if (e < ex_t.size() && c < cash_t.size()) {
                    const int k = total_cmp(ex_t[e], cash_t[c]);
                    if (k < 0) {
                        next = ex_t[e++];
                    } else if (k == 0) {
                        next = ex_t[e++];
                        ++c;
                    } else {
                        next = cash_t[c++];
                    }
                } else if (e < ex_t.size()) {
                    next = ex_t[e++];
                } else {
                    next = cash_t[c++];
                }
                if (next < T) times.push_back(next);
            }
            times.push_back(T);
        } else {
            const uint32_t base = rq->mesh_kind == FOHIP_MESH_CASH_DIVIDEND_DATES ? 1u : rq->base_intervals;
            int64_t node = 0;
            size_t event = 0;
            while (node >= 0 || event < cash_t.size()) {
                const bool have_base = node >= 0;
                double bt = 0.0;
                if (have_base)
                    bt = static_cast<uint32_t>(node) == base
                             ? T
                             : T * (static_cast<double>(static_cast<uint32_t>(node)) / static_cast<double>(base));
                auto advance = [&]() { node = static_cast<uint32_t>(node) == base ? -1 : node + 1; };
                if (have_base && event < cash_t.size()) {
                    const int k = total_cmp(bt, cash_t[event]);
                    if (k < 0) {
                        times.push_back(bt);
                        advance();
                    } else if (k == 0) {
                        times.push_back(bt);
                        advance();
                        ++event;
                    } else {
                        times.push_back(cash_t[event++]);
                    }
                } else if (have_base) {
                    times.push_back(bt);
                    advance();
                } else {
                    times.push_back(cash_t[event++]);
                }
            }
        }
        const size_t group_row = rq->group_offsets[g];
        {
            int rc = fohip_validate_time_grid(times.data(), times.size(), err);
            if (rc) {
                err->row = static_cast<int64_t>(group_row);
                return rc;
            }
            if (times.size() < 2) return fail(err, FOHIP_E_EXERCISE_MESH, 0, "row %zu mesh has no interval", group_row);
        }
        // build_row_curves (pricing.rs:862-891)
        const double yield = has_cash_schedule ? 0.0 : rq->dividend_yields[g];
        p.logd.resize(times.size());
        p.q.resize(times.size());
        for (size_t i = 0; i < times.size(); ++i) {
            p.logd[i] = i == 0 ? 0.0 : -rate * times[i];
            p.q[i] = i == 0 ? 0.0 : yield * times[i];
        }
        {
            int rc = fohip_validate_log_discount_curve(times.data(), p.logd.data(), times.size(), err);
            if (!rc) rc = fohip_validate_cumulative_yield_curve(times.data(), p.q.data(), times.size(), err);
            if (rc) {
                err->row = static_cast<int64_t>(group_row);
                return rc;
            }
        }
        // build_row_schedule (pricing.rs:893-965)
        p.cash.assign(times.size(), 0.0);
        p.has_cash.assign(times.size(), 0);
        std::vector<uint32_t> cash_node(cash_t.size());
        for (size_t e = 0; e < cash_t.size(); ++e) {
            const int64_t idx = search_time(times, cash_t[e]);
            if (idx < 0) return fail(err, FOHIP_E_CASH_SCHEDULE, 0, "row %zu cash event %zu has no grid node", group_row, e);
            cash_node[e] = static_cast<uint32_t>(idx);
            if (escrowed) continue;   // the risky prepaid forward carries no cash events
            p.cash[idx] = cash_a[e];
            p.has_cash[idx] = 1;
        }
        if (escrowed) {
            int rc = escrowed_plan(p, times, cash_node, cash_a, rq->initial_spots[g], rq->volatilities[g],
                                   rq->escrowed_volatility, group_row, err);
            if (rc) return rc;
        }
        p.obs_nodes.clear();
        if (rq->mesh_kind == FOHIP_MESH_EXPLICIT) {
            if (rq->origin_exercise == FOHIP_ORIGIN_INCLUDED) p.obs_nodes.push_back(0);
            for (size_t e = 0; e < ex_t.size(); ++e) {
                const int64_t idx = search_time(times, ex_t[e]);
                if (idx < 0) return fail(err, FOHIP_E_EXERCISE_MESH, 0, "row %zu exercise event %zu has no grid node", group_row, e);
                p.obs_nodes.push_back(static_cast<uint32_t>(idx));
            }
        } else {
            const uint32_t start =
                (rq->mesh_kind == FOHIP_MESH_CASH_DIVIDEND_DATES || rq->origin_exercise == FOHIP_ORIGIN_EXCLUDED) ? 1u : 0u;
            for (uint32_t n = start; n + 1 < times.size(); ++n) p.obs_nodes.push_back(n);
        }
        p.obs_nodes.push_back(static_cast<uint32_t>(times.size() - 1));
        p.S0 = rq->initial_spots[g];
        if (!escrowed) p.sigma = rq->volatilities[g];   // escrowed: sigma_F, set by escrowed_plan
        p.policy = rq->dividend_model == FOHIP_DIVIDENDS_CASH ? rq->cash_policy : 0;
        p.strike_begin = rq->group_offsets[g];
        p.strike_end = rq->group_offsets[g + 1];
        p.nt = nt_of(g);
        p.ne = ne_of(g);
        p.train_seed = rq->training_seed;
        p.eval_seed = rq->evaluation_seed;
        p.train_domain = rq->training_domain_start + static_cast<uint32_t>(g);
        p.eval_domain = rq->evaluation_domain_start + static_cast<uint32_t>(g);
        return FOHIP_OK;
    };
    const size_t hw = std::max<size_t>(1, std::thread::hardware_concurrency());
    const size_t nthreads = groups >= 1024 ? std::min<size_t>({hw, 16, groups / 256}) : 1;
    if (nthreads <= 1) {
        for (size_t g = 0; g < groups; ++g) {
            const int rc = build_one(g, err);
            if (rc) return rc;
        }
        return FOHIP_OK;
    }
    std::vector<size_t> first_failure(nthreads, groups);
    {
        std::vector<std::thread> pool;
        pool.reserve(nthreads);
        for (size_t k = 0; k < nthreads; ++k) {
            pool.emplace_back([&, k]() {
                fohip_error local;
                std::memset(&local, 0, sizeof local);
                const size_t g0 = groups * k / nthreads, g1 = groups * (k + 1) / nthreads;
                for (size_t g = g0; g < g1; ++g)
                    if (build_one(g, &local)) {
                        first_failure[k] = g;
                        return;
                    }
            });
        }
        for (std::thread &t : pool) t.join();
    }
    for (size_t k = 0; k < nthreads; ++k)
        if (first_failure[k] < groups) return build_one(first_failure[k], err);
    return FOHIP_OK;
}

extern "C" int fohip_price_lsm_grouped(fohip_executor *ex, const fohip_grouped_request *rq, fohip_batch_result *res,
                                       fohip_error *err) {
    clear_error(err);
    if (!ex || !rq || !res) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null executor, request or result");
    std::lock_guard<std::mutex> lock(ex->mu);
    FitParams fp;
    int rc = parse_regression(rq->polynomial_degree, rq->rank_tolerance, rq->condition_limit, rq->tie_rule,
                              rq->insufficient_support, fp, err);
    if (rc) return rc;
    std::vector<GroupPlan> plans;
    std::vector<double> strikes;
    rc = build_grouped_plans(rq, plans, strikes, err);
    if (rc) return rc;
    RunSpec spec;
    spec.fp = fp;
    spec.greeks = rq->greeks;
    set_method(spec, ex->precision);
    std::vector<RowOut> out;
    rc = run_groups(ex, plans, strikes, spec, out, nullptr, nullptr, err);
    if (rc) return rc;
    std::vector<uint64_t> row_paths(out.size());
    for (const auto &p : plans)
        for (uint32_t r = p.strike_begin; r < p.strike_end; ++r) row_paths[r] = p.ne;
    for (size_t r = 0; r < out.size(); ++r) {
        const RowOut &o = out[r];
        if (o.sum.early + static_cast<uint64_t>(o.sum.expiry) != row_paths[r]) {
            set_fields(err, static_cast<int64_t>(r), -1, -1, -1, 0, 0);
            return fail(err, FOHIP_E_STOPPING_COUNT_MISMATCH, FOHIP_PHASE_EVALUATION,
                        "row %zu stopping counts %llu do not partition %llu evaluation paths", r,
                        (unsigned long long)(o.sum.early + static_cast<uint64_t>(o.sum.expiry)),
                        (unsigned long long)row_paths[r]);
        }
    }
    for (size_t r = 0; r < out.size(); ++r) {
        const RowOut &o = out[r];
        if (res->prices) res->prices[r] = o.price;
        if (res->sample_standard_deviations) res->sample_standard_deviations[r] = o.sd;
        if (res->standard_errors) res->standard_errors[r] = o.se;
        if (res->origin_decisions_fitted) res->origin_decisions_fitted[r] = o.sum.origin_fitted;
        if (res->origin_exercised) res->origin_exercised[r] = o.sum.origin_exercised;
        if (res->regression_nodes) res->regression_nodes[r] = o.sum.regression_nodes;
        if (res->no_training_support_nodes) res->no_training_support_nodes[r] = o.sum.no_support_nodes;
        if (res->insufficient_exercise_support_nodes) res->insufficient_exercise_support_nodes[r] = o.sum.insufficient_nodes;
        if (res->minimum_in_the_money_training_paths) res->minimum_in_the_money_training_paths[r] = o.sum.min_itm;
        if (res->maximum_regression_condition_numbers) res->maximum_regression_condition_numbers[r] = o.sum.max_condition;
        if (res->early_exercise_paths) res->early_exercise_paths[r] = o.sum.early;
        if (res->expiry_paths) res->expiry_paths[r] = o.sum.expiry;
        if (res->near_tie_training_decisions) res->near_tie_training_decisions[r] = o.ties_training;
        if (res->near_tie_evaluation_decisions) res->near_tie_evaluation_decisions[r] = o.ties_evaluation;
        double *cols[4] = {res->delta, res->vega, res->rho, res->dividend_yield};
        double *secols[4] = {res->delta_se, res->vega_se, res->rho_se, res->dividend_yield_se};
        for (int a = 0; a < 4; ++a) {
            if (cols[a]) cols[a][r] = (rq->greeks & (1u << a)) ? o.greek[a] : std::nan("");
            if (secols[a]) secols[a][r] = (rq->greeks & (1u << a)) ? o.greek_se[a] : std::nan("");
        }
    }
    // keep the per-group layout of the last call for fohip_grouped_diagnostics
    ex->last_obs.clear();
    ex->last_strike_begin.clear();
    ex->last_strike_end.clear();
    for (const auto &p : plans) {
        ex->last_obs.push_back(static_cast<uint32_t>(p.obs_nodes.size()));
        ex->last_strike_begin.push_back(p.strike_begin);
        ex->last_strike_end.push_back(p.strike_end);
    }
    return FOHIP_OK;
}

extern "C" int fohip_grouped_diagnostics(fohip_executor *ex, const fohip_grouped_request *rq, uint32_t group,
                                         fohip_node_diagnostic *out, size_t capacity, uint32_t *exercise_rows,
                                         fohip_error *err) {
    clear_error(err);
    if (!ex || !rq || !out) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    std::lock_guard<std::mutex> lock(ex->mu);
    FitParams fp;
    int rc = parse_regression(rq->polynomial_degree, rq->rank_tolerance, rq->condition_limit, rq->tie_rule,
                              rq->insufficient_support, fp, err);
    if (rc) return rc;
    std::vector<GroupPlan> plans;
    std::vector<double> strikes;
    rc = build_grouped_plans(rq, plans, strikes, err);
    if (rc) return rc;
    if (group >= plans.size()) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "group %u out of range", group);
    // rerun the single group (identical domains and arithmetic) and return its fitted rules
    GroupPlan p = plans[group];
    const uint32_t ns = p.strike_end - p.strike_begin;
    std::vector<double> gstrikes(strikes.begin() + p.strike_begin, strikes.begin() + p.strike_end);
    p.strike_begin = 0;
    p.strike_end = ns;
    RunSpec spec;
    spec.fp = fp;
    spec.greeks = 0;
    set_method(spec, ex->precision);
    std::vector<RowOut> o;
    std::vector<Model> models;
    std::vector<GroupPlan> one{p};
    rc = run_groups(ex, one, gstrikes, spec, o, &models, nullptr, err);
    if (rc) return rc;
    const uint32_t obs = static_cast<uint32_t>(p.obs_nodes.size());
    const size_t need = static_cast<size_t>(ns) * (obs - 1);
    if (exercise_rows) *exercise_rows = obs - 1;
    if (capacity < need) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "diagnostics capacity %zu < %zu", capacity, need);
    for (uint32_t s = 0; s < ns; ++s) {
        for (uint32_t r = 0; r + 1 < obs; ++r) {
            const Model &m = models[static_cast<size_t>(s) * obs + r];
            fohip_node_diagnostic &d = out[static_cast<size_t>(s) * (obs - 1) + r];
            std::memset(&d, 0, sizeof d);
            d.node = p.obs_nodes[r];
            d.kind = m.kind == kRuleOrigin ? 0 : m.kind == kRuleNoSupport ? 1 : m.kind == kRuleContinue ? 2 : 3;
            d.exercise = m.exercise;
            d.estimated_continuation = m.estimate;
            d.in_the_money_paths = m.itm;
            d.observations = m.itm;
            d.columns = static_cast<uint32_t>(fp.degree + 1);
            d.rank = static_cast<uint32_t>(m.rank);
            d.condition_number = m.condition;
            d.predictor_mean = m.mean;
            d.predictor_scale = m.scale;
            for (int k = 0; k < 4; ++k) {
                d.alpha[k] = m.alpha[k];
                d.beta[k] = m.beta[k];
            }
            for (int k = 0; k < 5; ++k) d.coefficients[k] = m.coef[k];
        }
    }
    return FOHIP_OK;
}

// ======================================================================== path export/probes

// GbmPathPlan::simulate_seeded (crates/finoptions-sim/src/executor.rs:34, seeded.rs:38-77) of one
// plan into `device_out` (caller device memory, binary32 when f32) or, when it is null, into the
// executor's path buffer copied to `host_out` as binary64. Work queued on `after` (a caller stream)
// completes first; the call returns once the states are complete.
static int simulate_export(fohip_executor *ex, const fohip_path_plan *pl, double S0, uint64_t paths, uint64_t seed,
                           uint32_t domain, bool f32, void *device_out, hipStream_t after, double *host_out,
                           fohip_error *err) {
    int rc = fohip_validate_event_schedule(pl->times, pl->nodes, pl->dividend_nodes, pl->dividend_amounts, pl->dividends,
                                           pl->exercise_nodes, pl->exercise_count, pl->expiry_node, err);
    if (!rc) rc = fohip_validate_log_discount_curve(pl->times, pl->log_discounts, pl->nodes, err);
    if (!rc) rc = fohip_validate_cumulative_yield_curve(pl->times, pl->cumulative_yields, pl->nodes, err);
    if (!rc) rc = fohip_validate_constant_volatility(pl->volatility, err);
    if (rc) return rc;
    if (!canonical_zero_or_positive(S0))
        return fail(err, FOHIP_E_INVALID_INITIAL_SPOT, 0, "initial spot must be finite, nonnegative, and canonical; observed %.17g", S0);
    if (paths == 0) return fail(err, FOHIP_E_EMPTY_PATH_SET, 0, "simulation path count must be positive");
    if (paths > 0xFFFFFFFFull) return fail(err, FOHIP_E_PATH_NAMESPACE_OVERFLOW, 0, "path count exceeds the 32-bit path namespace");
    HIP_TRY(hipSetDevice(ex->device));
    GroupPlan p;
    p.times.assign(pl->times, pl->times + pl->nodes);
    p.logd.assign(pl->log_discounts, pl->log_discounts + pl->nodes);
    p.q.assign(pl->cumulative_yields, pl->cumulative_yields + pl->nodes);
    p.cash.assign(pl->nodes, 0.0);
    p.has_cash.assign(pl->nodes, 0);
    for (size_t i = 0; i < pl->dividends; ++i) {
        p.cash[pl->dividend_nodes[i]] = pl->dividend_amounts[i];
        p.has_cash[pl->dividend_nodes[i]] = 1;
    }
    p.obs_nodes.assign(pl->exercise_nodes, pl->exercise_nodes + pl->exercise_count);
    p.obs_nodes.push_back(pl->expiry_node);
    p.S0 = S0;
    p.sigma = pl->volatility;
    p.policy = pl->dividend_policy;
    p.strike_begin = 0;
    p.strike_end = 0;
    p.nt = paths;
    p.ne = 0;
    p.train_seed = seed;
    p.train_domain = domain;
    std::vector<GroupPlan> plans{p};
    std::vector<double> none;
    GroupTables gt;
    uint64_t path_elems, value_elems, model_count, part_t, part_e, count_slots;
    int64_t bad_group, bad_node;
    build_tables(plans, none, false, gt, path_elems, value_elems, model_count, part_t, part_e, count_slots, err,
                 bad_group, bad_node);
    const size_t real = f32 ? 4 : 8;
    if (device_out) {
        // the destination must be device memory of this executor's device holding the whole matrix
        hipPointerAttribute_t attr;
        void *base = nullptr;
        size_t range = 0;
        const bool ok = hipPointerGetAttributes(&attr, device_out) == hipSuccess && attr.type == hipMemoryTypeDevice &&
                        attr.device == ex->device &&
                        hipMemGetAddressRange(&base, &range, device_out) == hipSuccess &&
                        static_cast<char *>(device_out) + path_elems * real <= static_cast<char *>(base) + range;
        if (!ok) {
            (void)hipGetLastError();
            return fail(err, FOHIP_E_INVALID_ARGUMENT, 0,
                        "device output must be device memory on device %d holding %llu elements", ex->device,
                        (unsigned long long)path_elems);
        }
    }
    const size_t tb_bytes[] = {sizeof(DevGroup), gt.log_drift.size() * 8, gt.diffusion.size() * 8, gt.log_drift32.size() * 4,
                               gt.diffusion32.size() * 4, gt.cash.size() * 8, gt.node_flags.size() * 4,
                               gt.steps.size() * sizeof(StepRec)};
    const void *tb_src[] = {gt.groups.data(), gt.log_drift.data(), gt.diffusion.data(), gt.log_drift32.data(),
                            gt.diffusion32.data(), gt.cash.data(), gt.node_flags.data(), gt.steps.data()};
    constexpr int kParts = 8;
    size_t total = 0;
    size_t offs[kParts];
    for (int i = 0; i < kParts; ++i) {
        offs[i] = total;
        total += (tb_bytes[i] + 255) & ~size_t(255);
    }
    HIP_TRY(ex->slot[0].tables.ensure(total));
    if (!device_out) HIP_TRY(ex->slot[0].paths.ensure(path_elems * real));
    void *dst = device_out ? device_out : ex->slot[0].paths.ptr;
    HIP_TRY(ex->slot[0].fail_t.ensure(8));
    std::vector<uint2> items;
    for (uint32_t c = 0; c < gt.groups[0].chunks_t; ++c) items.push_back(make_uint2(0, c));
    HIP_TRY(ex->items_t.ensure(items.size() * sizeof(uint2)));
    char *tb = ex->slot[0].tables.as<char>();
    hipStream_t st = ex->stream;
    if (after && after != st) {
        hipEvent_t ready;
        HIP_TRY(hipEventCreateWithFlags(&ready, hipEventDisableTiming));
        hipError_t e = hipEventRecord(ready, after);
        if (e == hipSuccess) e = hipStreamWaitEvent(st, ready, 0);
        (void)hipEventDestroy(ready);
        HIP_TRY(e);
    }
    for (int i = 0; i < kParts; ++i) HIP_TRY(hipMemcpyAsync(tb + offs[i], tb_src[i], tb_bytes[i], hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(ex->items_t.ptr, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemsetAsync(ex->slot[0].fail_t.ptr, 0xFF, 8, st));
    DevTables t;
    std::memset(&t, 0, sizeof t);
    t.groups = reinterpret_cast<const DevGroup *>(tb + offs[0]);
    t.log_drift = reinterpret_cast<const double *>(tb + offs[1]);
    t.diffusion = reinterpret_cast<const double *>(tb + offs[2]);
    t.log_drift32 = reinterpret_cast<const float *>(tb + offs[3]);
    t.diffusion32 = reinterpret_cast<const float *>(tb + offs[4]);
    t.cash = reinterpret_cast<const double *>(tb + offs[5]);
    t.node_flags = reinterpret_cast<const uint32_t *>(tb + offs[6]);
    t.steps = reinterpret_cast<const StepRec *>(tb + offs[7]);
    HIP_TRY(launch_simulate(st, f32, t, ex->items_t.as<uint2>(), static_cast<uint32_t>(items.size()), dst, 0,
                            ex->slot[0].fail_t.as<unsigned long long>()));
    unsigned long long failed = 0;
    HIP_TRY(hipMemcpyAsync(&failed, ex->slot[0].fail_t.ptr, 8, hipMemcpyDeviceToHost, st));
    if (device_out) {
        HIP_TRY(hipStreamSynchronize(st));
    } else if (f32) {
        std::vector<float> tmp(path_elems);
        HIP_TRY(hipMemcpyAsync(tmp.data(), dst, path_elems * 4, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
        for (uint64_t i = 0; i < path_elems; ++i) host_out[i] = tmp[i];
    } else {
        HIP_TRY(hipMemcpyAsync(host_out, dst, path_elems * 8, hipMemcpyDeviceToHost, st));
        HIP_TRY(hipStreamSynchronize(st));
    }
    if (failed != kNoError) {
        set_fields(err, -1, -1, static_cast<int64_t>(failed), -1, 0, 0);
        return fail(err, FOHIP_E_NON_FINITE_TRANSITION, FOHIP_PHASE_TRAINING_SIMULATION,
                    "GBM path transition at path %llu produced a non-finite spot", failed);
    }
    return FOHIP_OK;
}

extern "C" int fohip_simulate_gbm_seeded(fohip_executor *ex, const fohip_path_plan *pl, double S0, uint64_t paths,
                                         uint64_t seed, uint32_t domain, double *out, fohip_error *err) {
    clear_error(err);
    if (!ex || !pl || !out) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    std::lock_guard<std::mutex> lock(ex->mu);
    return simulate_export(ex, pl, S0, paths, seed, domain, ex->precision == FOHIP_PRECISION_F32_BULK, nullptr,
                           nullptr, out, err);
}

extern "C" int fohip_simulate_gbm_seeded_device(fohip_executor *ex, const fohip_path_plan *pl, double S0,
                                                uint64_t paths, uint64_t seed, uint32_t domain, void *device_out,
                                                int32_t dtype, void *stream, fohip_error *err) {
    clear_error(err);
    if (!ex || !pl || !device_out) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    if (dtype != FOHIP_DTYPE_F64 && dtype != FOHIP_DTYPE_F32)
        return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "unknown output dtype %d", dtype);
    std::lock_guard<std::mutex> lock(ex->mu);
    return simulate_export(ex, pl, S0, paths, seed, domain, dtype == FOHIP_DTYPE_F32, device_out,
                           static_cast<hipStream_t>(stream), nullptr, err);
}

extern "C" int fohip_philox_blocks(fohip_executor *ex, uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths,
                                   uint32_t blocks, uint32_t *out, fohip_error *err) {
    clear_error(err);
    if (!ex || !out) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    std::lock_guard<std::mutex> lock(ex->mu);
    HIP_TRY(hipSetDevice(ex->device));
    const size_t bytes = static_cast<size_t>(paths) * blocks * 16;
    if (bytes == 0) return FOHIP_OK;
    HIP_TRY(ex->probe_a.ensure(bytes));
    HIP_TRY(launch_philox_blocks(ex->stream, seed, domain, path0, paths, blocks, ex->probe_a.as<uint32_t>()));
    HIP_TRY(hipMemcpyAsync(out, ex->probe_a.ptr, bytes, hipMemcpyDeviceToHost, ex->stream));
    HIP_TRY(hipStreamSynchronize(ex->stream));
    return FOHIP_OK;
}

extern "C" int fohip_gbm_exp_probe(fohip_executor *ex, const double *x, size_t n, double *out, fohip_error *err) {
    clear_error(err);
    if (!ex || (n && (!x || !out))) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    if (n > 0xFFFFFFFFu) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "at most 2^32 - 1 arguments per call");
    if (n == 0) return FOHIP_OK;
    std::lock_guard<std::mutex> lock(ex->mu);
    HIP_TRY(hipSetDevice(ex->device));
    HIP_TRY(ex->probe_a.ensure(n * 8));
    HIP_TRY(ex->probe_b.ensure(n * 8));
    HIP_TRY(hipMemcpyAsync(ex->probe_a.ptr, x, n * 8, hipMemcpyHostToDevice, ex->stream));
    HIP_TRY(launch_gbm_exp(ex->stream, ex->probe_a.as<double>(), ex->probe_b.as<double>(), static_cast<uint32_t>(n)));
    HIP_TRY(hipMemcpyAsync(out, ex->probe_b.ptr, n * 8, hipMemcpyDeviceToHost, ex->stream));
    HIP_TRY(hipStreamSynchronize(ex->stream));
    return FOHIP_OK;
}

extern "C" int fohip_stream_normals(fohip_executor *ex, uint64_t seed, uint32_t domain, uint32_t path0, uint32_t paths,
                                    uint32_t count, double *out, uint32_t *draws, fohip_error *err) {
    clear_error(err);
    if (!ex || !out || !draws) return fail(err, FOHIP_E_INVALID_ARGUMENT, 0, "null argument");
    std::lock_guard<std::mutex> lock(ex->mu);
    HIP_TRY(hipSetDevice(ex->device));
    const size_t n = static_cast<size_t>(paths) * count;
    if (n == 0) return FOHIP_OK;
    HIP_TRY(ex->probe_a.ensure(n * 8));
    HIP_TRY(ex->probe_b.ensure(n * 4));
    HIP_TRY(ex->probe_c.ensure(sizeof(int)));
    HIP_TRY(hipMemsetAsync(ex->probe_c.ptr, 0, sizeof(int), ex->stream));
    HIP_TRY(launch_stream_normals(ex->stream, seed, domain, path0, paths, count, ex->probe_a.as<double>(),
                                  ex->probe_b.as<uint32_t>(), ex->probe_c.as<int>()));
    int failed = 0;
    HIP_TRY(hipMemcpyAsync(out, ex->probe_a.ptr, n * 8, hipMemcpyDeviceToHost, ex->stream));
    HIP_TRY(hipMemcpyAsync(draws, ex->probe_b.ptr, n * 4, hipMemcpyDeviceToHost, ex->stream));
    HIP_TRY(hipMemcpyAsync(&failed, ex->probe_c.ptr, sizeof(int), hipMemcpyDeviceToHost, ex->stream));
    HIP_TRY(hipStreamSynchronize(ex->stream));
    if (failed) return fail(err, FOHIP_E_NORMAL_REJECTION_LIMIT, 0, "random-normal rejection limit reached");
    return FOHIP_OK;
}
