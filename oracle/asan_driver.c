/* Sanitizer driver for the CPU oracle (test infrastructure only, never shipped).
 *
 * Built with -fsanitize=address,undefined together with lsm_oracle.c (oracle/Makefile, target
 * asan-check) and run by tests/test_oracle_asan.py: it exercises the oracle's entry points on
 * small requests of every dividend model, mesh kind and the ragged path counts, serial and
 * group-parallel, and checks that they succeed and agree with each other. A sanitizer report
 * makes the run fail. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "lsm_oracle.h"

static int failures = 0;
#define CHECK(cond, what)                                                  \
    do {                                                                   \
        if (!(cond)) {                                                     \
            fprintf(stderr, "check failed: %s (%s:%d)\n", what, __FILE__, __LINE__); \
            ++failures;                                                    \
        }                                                                  \
    } while (0)

typedef struct {
    double prices[16], sds[16], ses[16], max_condition[16];
    uint8_t origin_fitted[16], origin_exercised[16];
    uint32_t regression_nodes[16], no_support_nodes[16], insufficient_nodes[16], min_itm_paths[16];
    uint32_t early_paths[16], expiry_paths[16];
    uint64_t ties_t[16], ties_e[16];
    or_batch_output out;
} batch_buffers;

static void bind(batch_buffers *b) {
    memset(b, 0, sizeof *b);
    b->out.prices = b->prices;
    b->out.sds = b->sds;
    b->out.ses = b->ses;
    b->out.origin_fitted = b->origin_fitted;
    b->out.origin_exercised = b->origin_exercised;
    b->out.regression_nodes = b->regression_nodes;
    b->out.no_support_nodes = b->no_support_nodes;
    b->out.insufficient_nodes = b->insufficient_nodes;
    b->out.min_itm_paths = b->min_itm_paths;
    b->out.max_condition = b->max_condition;
    b->out.early_paths = b->early_paths;
    b->out.expiry_paths = b->expiry_paths;
    b->out.near_tie_training = b->ties_t;
    b->out.near_tie_evaluation = b->ties_e;
}

int main(void) {
    /* three groups: 2 + 3 + 1 strikes */
    static const uint32_t offsets[4] = {0, 2, 5, 6};
    static const double spots[3] = {100.0, 95.0, 110.0}, mats[3] = {1.0, 0.75, 1.5};
    static const double rates[3] = {0.03, 0.01, 0.05}, vols[3] = {0.25, 0.4, 0.2}, yields[3] = {0.02, 0.0, 0.01};
    static const double strikes[6] = {90.0, 110.0, 80.0, 95.0, 120.0, 100.0};
    static const int64_t cash_off[4] = {0, 1, 3, 3};
    static const double cash_t[3] = {0.5, 0.2, 0.6}, cash_a[3] = {1.5, 0.5, 2.0};
    static const int64_t ex_off[4] = {0, 3, 5, 9};
    static const double ex_t[9] = {0.25, 0.5, 0.75, 0.3, 0.6, 0.2, 0.5, 0.9, 1.2};
    static const uint32_t nt[3] = {2048, 1024, 3000}, ne[3] = {1500, 2048, 1024};
    or_grouped_input in;
    memset(&in, 0, sizeof in);
    in.groups = 3;
    in.group_offsets = offsets;
    in.initial_spots = spots;
    in.maturities = mats;
    in.rates = rates;
    in.volatilities = vols;
    in.rows = 6;
    in.strikes = strikes;
    in.training_paths = 4096;
    in.evaluation_paths = 4096;
    in.training_seed = 11;
    in.evaluation_seed = 11;
    in.training_domain_start = 1;
    in.evaluation_domain_start = 10;
    in.degree = 4;
    in.rank_tolerance = 1e-12;
    in.condition_limit = 1e10;
    in.insufficient_support = 1;
    in.base_intervals = 12;

    batch_buffers serial, parallel;
    or_error err;
    for (int model = 0; model < 3; ++model) {
        for (int mesh = 0; mesh < 3; ++mesh) {
            if (mesh == 1 && model == 0) continue;   /* the cash-date mesh needs cash dividends */
            in.dividend_model = model;
            in.dividend_yields = model == 0 ? yields : NULL;
            in.cash_offsets = model ? cash_off : NULL;
            in.cash_times = model ? cash_t : NULL;
            in.cash_amounts = model ? cash_a : NULL;
            in.mesh_kind = mesh;
            in.exercise_offsets = mesh == 2 ? ex_off : NULL;
            in.exercise_times = mesh == 2 ? ex_t : NULL;
            in.training_path_counts = mesh == 2 ? nt : NULL;
            in.evaluation_path_counts = mesh == 2 ? ne : NULL;
            bind(&serial);
            bind(&parallel);
            int rc = or_price_grouped(&in, &serial.out, 1, &err);
            CHECK(rc == 0, "serial grouped pricing");
            rc = or_price_grouped(&in, &parallel.out, 3, &err);
            CHECK(rc == 0, "group-parallel grouped pricing");
            CHECK(memcmp(serial.prices, parallel.prices, sizeof serial.prices) == 0, "parallel == serial");
            for (uint32_t g = 0; g < 3; ++g) {
                double d[3], v[3], r[3], q[3];
                rc = or_grouped_pathwise_greeks(&in, g, d, v, r, q, &err);
                CHECK(rc == 0, "pathwise greeks");
                static double paths[64 * 4096];
                rc = or_grouped_paths(&in, g, 1, paths, &err);
                CHECK(rc == 0, "path export");
                double times[64];
                uint32_t exn[64], exc = 0, dn[64], dc = 0;
                double da[64];
                int nodes = or_group_grid(&in, g, times, 64, exn, &exc, dn, da, &dc);
                CHECK(nodes > 1, "group grid");
            }
            printf("model %d mesh %d: prices %.6f %.6f %.6f\n", model, mesh, serial.prices[0], serial.prices[2],
                   serial.prices[5]);
        }
    }
    /* streams and normals */
    uint64_t bits[64];
    or_stream_u64(5, 3, 7, 64, bits);
    double normals[32];
    uint32_t draws[32];
    CHECK(or_stream_normals(5, 3, 7, 32, normals, draws) == 0, "stream normals");
    printf("sanitized oracle run: %d failed checks\n", failures);
    return failures ? 1 : 0;
}
