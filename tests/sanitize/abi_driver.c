/*
 * Sanitizer driver for the C ABI's host code, linked against the engine
 * library built with host-side -fsanitize=address,undefined by
 * tests/test_sanitize.py.  Without a GPU every entry point must reject bad
 * arguments, or report the missing device, with an error code and a
 * message; the sanitizers abort on any memory error or undefined behaviour
 * in the host paths they run (argument parsing, option validation,
 * workspace sizing, object set-up and tear-down).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

#include "winds_bs.h"
#include "winds_ground.h"
#include "winds_march.h"
#include "winds_mlp.h"
#include "winds_surrogate.h"

static int fails = 0;

static void expect_error(int code, const char *what) {
    const char *msg = winds_last_error();
    if (code == WINDS_OK || !msg || !msg[0]) {
        fprintf(stderr, "%s: code %d, message '%s'\n", what, code, msg ? msg : "(null)");
        fails++;
    }
}

static int call(const char *type, const char *cmod, double d, const int64_t fd[4],
                const int64_t pd[4], const double hw[3]) {
    static double F[64], P[64], g[64], u[64], L[64];
    memset(F, 0, sizeof F);
    return winds_bs_eval(F, F, fd, P, pd, g, g, d, cmod, 1000.0, type, "true", hw, u, L);
}

int main(void) {
    winds_bs_opts o;
    winds_bs_opts_default(&o);
    if (!winds_version() || !strstr(winds_version(), "gfx950")) fails++;

    /* workspace sizing over every option and a range of sizes */
    const int64_t sizes[] = {0, 1, 63, 64, 65, 4096, 133920, 1 << 20};
    for (int prec = 0; prec < 2; prec++)
        for (int core = 0; core < 2; core++)
            for (int sem = 0; sem < 2; sem++)
                for (int n = 1; n <= 4; n++)
                    for (int i = 0; i < 8; i++)
                        for (int j = 0; j < 8; j++) {
                            winds_bs_opts q = o;
                            q.precision = prec;
                            q.core = core;
                            q.semantics = sem;
                            q.vat_n = n;
                            (void)winds_bs_workspace_bytes(sizes[i], sizes[j], &q);
                        }
    winds_bs_opts bad = o;
    bad.guard_kappa = NAN;
    bad.precision = WINDS_PREC_FP32C;
    if (winds_bs_workspace_bytes(10, 10, &bad) != 0) fails++;
    bad = o;
    bad.core = 7;
    if (winds_bs_workspace_bytes(10, 10, &bad) != 0) fails++;
    if (winds_bs_workspace_bytes(10, 10, NULL) != 0) fails++;

    /* the host-array entry point: bad strings, dimensions, hardware */
    const int64_t fd[4] = {2, 3, 2, 1}, pd[4] = {2, 3, 2, 1};
    const int64_t fbad[4] = {2, 2, 2, 1}, pbad[4] = {2, 4, 2, 1};
    const double hw[3] = {256, 1, 0}, hw_t[3] = {100, 1, 0}, hw_g[3] = {256, 2, 0};
    expect_error(call("xxxx", "none", 1e-4, fd, pd, hw), "unknown type");
    expect_error(call("fuln", "bogus", 1e-4, fd, pd, hw), "unknown cmod");
    expect_error(call("fuln", "visc", 2.5, fd, pd, hw), "fractional Vatistas order");
    expect_error(call("fuln", "none", 1e-4, fbad, pd, hw), "bad F dims");
    expect_error(call("fuln", "none", 1e-4, fd, pbad, hw), "bad P dims");
    expect_error(call("fuln", "none", 1e-4, fd, pd, hw_t), "bad thread count");
    expect_error(call("fuln", "none", 1e-4, fd, pd, hw_g), "two GPUs");
    expect_error(call("full", "none", 1e-4, fd, pd, hw), "no device");
    expect_error(call("fulc", "visc", 2, fd, pd, NULL), "no device, FP32C");
    expect_error(call("leng", "none", 1e-4, fd, pd, hw), "no device, lengths");
    expect_error(winds_bs_eval(NULL, NULL, fd, NULL, pd, NULL, NULL, 1e-4, "none", 1.0, "fuln",
                               "true", hw, NULL, NULL), "null arrays");
    expect_error(winds_bs_eval(NULL, NULL, NULL, NULL, NULL, NULL, NULL, 1e-4, NULL, 1.0, NULL,
                               NULL, NULL, NULL, NULL), "null everything");

    /* the device entry points: options, sizes, pointers */
    double t3[3] = {0, 0, 0}, s8[8] = {0, 0, 0, 1, 0, 0, 1, 0}, u3[3];
    char ws[4096];
    expect_error(winds_bs_eval_dev(t3, -1, s8, 1, &o, u3, ws, sizeof ws, NULL), "negative size");
    expect_error(winds_bs_eval_dev(NULL, 1, s8, 1, &o, u3, ws, sizeof ws, NULL), "null targets");
    expect_error(winds_bs_eval_dev(t3, 1, NULL, 1, &o, u3, ws, sizeof ws, NULL), "null segments");
    expect_error(winds_bs_eval_dev(t3, 1, s8, 1, &bad, u3, ws, sizeof ws, NULL), "bad options");
    expect_error(winds_bs_eval_dev_noalloc(t3, 1, s8, 1, &bad, u3, NULL), "noalloc bad options");
    int64_t cnt = 0;
    winds_bs_opts o32 = o;
    o32.precision = WINDS_PREC_FP32C;
    expect_error(winds_bs_guard_count(NULL, 1, 1, &o32, &cnt, NULL), "guard count, null workspace");
    expect_error(winds_bs_guard_count(ws, 1, 1, &o32, NULL, NULL), "guard count, null count");
    /* FP64 has no guard: zero re-evaluated targets, whatever the workspace */
    if (winds_bs_guard_count(NULL, 1, 1, &o, &cnt, NULL) != WINDS_OK || cnt != 0) fails++;

    /* batches */
    winds_batch *b = NULL;
    int64_t nt[2] = {5, -1}, ns[2] = {3, 3};
    expect_error(winds_batch_create(2, nt, ns, &o, NULL, &b), "negative batch size");
    expect_error(winds_batch_create(-1, nt, ns, &o, NULL, &b), "negative wake count");
    expect_error(winds_batch_create(1, NULL, ns, &o, NULL, &b), "null sizes");
    winds_batch_destroy(NULL);

    /* march, network, ground plane, surrogate geometry */
    winds_march *m = NULL;
    winds_march_cfg mc;
    memset(&mc, 0, sizeof mc);
    expect_error(winds_march_create(NULL, &m), "null march cfg");
    expect_error(winds_march_create(&mc, &m), "empty march cfg");
    mc.nb = 3;
    mc.nst = 31;
    mc.max_ages = 1440;
    mc.n_times = 100;
    expect_error(winds_march_create(&mc, &m), "march cfg without tables");
    winds_march_destroy(NULL);
    winds_mlp *net = NULL;
    expect_error(winds_mlp_create(0, 256, 4, 3, 1024, 1, &net), "network without inputs");
    expect_error(winds_mlp_create(49, 100, 4, 3, 1024, 1, &net), "hidden width not 32k");
    expect_error(winds_mlp_create(49, 256, 4, 3, 1024, 1, NULL), "null network out");
    winds_mlp_destroy(NULL);
    if (winds_mlp_num_params(NULL) != 0) fails++;
    winds_ground *gp = NULL;
    expect_error(winds_ground_create(0, 0, 1, 0, 1, 0.05, &gp), "no panels");
    expect_error(winds_ground_create(200, 0, 1, 0, 1, 0.05, &gp), "too many panels");
    expect_error(winds_ground_create(10, 1, 0, 0, 1, 0.05, &gp), "empty extent");
    expect_error(winds_ground_create(10, 0, 1, 0, 1, -1.0, &gp), "negative cut-off");
    expect_error(winds_ground_create(10, 0, 1, 0, 1, 0.05, &gp), "no device, ground");
    winds_ground_destroy(NULL);
    if (winds_ground_panels(NULL) != 0) fails++;
    expect_error(winds_ground_induced_dev(NULL, NULL, NULL, 1, NULL, NULL), "ground null");

    if (fails) {
        fprintf(stderr, "abi driver: %d check(s) failed\n", fails);
        return 1;
    }
    printf("abi driver: ok\n");
    return 0;
}
