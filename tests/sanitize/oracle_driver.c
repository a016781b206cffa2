/*
 * Sanitizer driver for the CPU oracle (oracle/bs_oracle.c), built with
 * -fsanitize=address,undefined by tests/test_sanitize.py.  Runs every oracle
 * entry point over seeded random wakes and the edge cases its callers meet:
 * zero-length segments, targets on segment end points and on the segment
 * line, empty target or segment sets, and non-square array shapes.  Exits
 * non-zero when a result is not finite where it must be, or when two forms
 * of the same sum disagree; the sanitizers abort on any memory or UB error.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

void oracle_bs_mex(const double *F1, const double *F2, const int64_t *fdims, const double *P,
                   const int64_t *pdims, const double *gamma, double d, double *uind,
                   int nthreads);
void oracle_bs_mex_range(const double *F1, const double *F2, const int64_t *fdims,
                         const double *P, const int64_t *pdims, const double *gamma, double d,
                         int64_t j0, int64_t j1, double *out3, int nthreads);
void oracle_bs_matlab(const double *F1, const double *F2, const int64_t *fdims, const double *P,
                      const int64_t *pdims, const double *gamma, const double *rc, double d,
                      int n, double co, double *uind, int nthreads);
void oracle_bs_quad(const double *F1, const double *F2, const int64_t *fdims, const double *P,
                    const int64_t *pdims, const double *gamma, const double *rc, double d,
                    int n, double co, double *uind);
void oracle_filament_length(const double *F1, const double *F2, const int64_t *fdims,
                            double *L);

static uint64_t rng = 0x9E3779B97F4A7C15ull;
static double urand(void) {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return (double)(rng >> 11) / 9007199254740992.0;
}

static int run(int64_t fns, int64_t fnt, int64_t fnb, int64_t pns, int64_t pnt, int64_t pnb,
               int edge) {
    const int64_t fd[4] = {fns, 3, fnt, fnb}, pd[4] = {pns, 3, pnt, pnb};
    const int64_t nf = fns * fnt * fnb, np = pns * pnt * pnb;
    double *F1 = malloc(sizeof(double) * (3 * nf + 1)), *F2 = malloc(sizeof(double) * (3 * nf + 1));
    double *P = malloc(sizeof(double) * (3 * np + 1)), *g = malloc(sizeof(double) * (nf + 1));
    double *rc = malloc(sizeof(double) * (nf + 1)), *L = malloc(sizeof(double) * (nf + 1));
    double *u1 = malloc(sizeof(double) * (3 * np + 1)), *u2 = malloc(sizeof(double) * (3 * np + 1));
    double *u3 = malloc(sizeof(double) * (3 * np + 1)), *u4 = malloc(sizeof(double) * (3 * np + 1));
    for (int64_t i = 0; i < 3 * nf; i++) {
        F1[i] = 4 * urand() - 2;
        F2[i] = F1[i] + 0.4 * urand() - 0.2;
    }
    for (int64_t i = 0; i < nf; i++) {
        g[i] = urand() - 0.5;
        rc[i] = 1e-3 + 1e-2 * urand();
    }
    for (int64_t i = 0; i < 3 * np; i++) P[i] = 4 * urand() - 2;
    if (edge && nf > 0 && np > 0) {
        /* a zero-length segment, a target on a start point, one mid-segment */
        for (int c = 0; c < 3; c++) F2[c * fns] = F1[c * fns];
        for (int c = 0; c < 3; c++) P[c * pns] = F1[c * fns + (nf > 1 ? 1 : 0)];
        if (np > 1 && nf > 1)
            for (int c = 0; c < 3; c++) P[c * pns + 1] = 0.5 * (F1[c * fns + 1] + F2[c * fns + 1]);
    }
    oracle_bs_mex(F1, F2, fd, P, pd, g, 1e-4, u1, 2);
    oracle_bs_mex_range(F1, F2, fd, P, pd, g, 1e-4, 0, np, u2, 2);
    oracle_bs_matlab(F1, F2, fd, P, pd, g, rc, 1e-4, 0, INFINITY, u3, 2);
    oracle_filament_length(F1, F2, fd, L);
    int bad = 0;
    for (int64_t j = 0; j < np; j++) {
        const int64_t m = j % pns, n = (j / pns) % pnt, p = j / (pns * pnt);
        for (int c = 0; c < 3; c++) {
            const double a = u1[m + pns * (c + 3 * (n + pnt * p))];
            const double b = u2[3 * j + c];
            /* the packed form is the same sum (NaN where the MEX arithmetic
               meets 0/0, e.g. a target on a zero-length segment) */
            if (!(a == b || (isnan(a) && isnan(b)))) bad |= 1;
        }
    }
    for (int n = 1; n <= 3; n++) {
        oracle_bs_matlab(F1, F2, fd, P, pd, g, rc, 0.0, n, 10.0, u3, 1);
        for (int64_t i = 0; i < 3 * np; i++)
            if (!isfinite(u3[i])) bad |= 2;   /* the Vatistas cores are finite everywhere */
    }
    if (np <= 64 && nf <= 64) {
        oracle_bs_quad(F1, F2, fd, P, pd, g, rc, 0.0, 2, INFINITY, u4);
        oracle_bs_matlab(F1, F2, fd, P, pd, g, rc, 0.0, 2, INFINITY, u3, 1);
        for (int64_t i = 0; i < 3 * np; i++)
            if (fabs(u4[i] - u3[i]) > 1e-9 * (1 + fabs(u4[i]))) bad |= 4;
    }
    for (int64_t i = 0; i < nf; i++)
        if (!(L[i] >= 0)) bad |= 8;
    free(F1); free(F2); free(P); free(g); free(rc); free(L); free(u1); free(u2); free(u3); free(u4);
    return bad;
}

int main(void) {
    int bad = 0;
    bad |= run(5, 7, 3, 4, 6, 2, 1);
    bad |= run(1, 1, 1, 1, 1, 1, 1);
    bad |= run(31, 40, 3, 31, 41, 3, 1);
    bad |= run(3, 2, 1, 0, 1, 1, 0);     /* no targets */
    bad |= run(0, 1, 1, 5, 1, 1, 0);     /* no segments */
    bad |= run(8, 8, 1, 8, 8, 1, 1);
    if (bad) {
        fprintf(stderr, "oracle driver: check failed (mask %d)\n", bad);
        return 1;
    }
    printf("oracle driver: ok\n");
    return 0;
}
