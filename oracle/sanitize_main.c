/*
 * sanitize_main.c -- TEST INFRASTRUCTURE ONLY: drives the oracle through every entry point on small
 * synthetic states, for builds with AddressSanitizer and UndefinedBehaviorSanitizer (oracle/Makefile
 * target `sanitize`).  Exits 0 when every call returned and the states stayed finite; the sanitizers
 * abort the run on the first invalid access or undefined operation.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "dyncore_oracle.h"

static unsigned long long lcg = 20261202ULL;
static double urand(void) { /* uniform in [-1, 1) */
    lcg = lcg * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(lcg >> 11) / 9007199254740992.0 * 2.0 - 1.0;
}

static int run(int nx, int ny, int nz, int moist, int raylevel, int nsteps) {
    orc_grid g = {nx, ny, nz, 1000.0, 250.0, 2.0, moist, raylevel, 5, 5};
    orc_state *s = orc_create(&g);
    if (!s) return 1;
    const size_t n = (size_t)nx * ny * nz;
    double *col = calloc((size_t)(nz > ny ? nz : ny), sizeof(double));
    double *f = malloc(n * sizeof(double));
    /* columns: a stable, moist-adiabatic-like profile on a uniform grid (initializer.cpp:840-928 shape) */
    for (int c = 0; c < 16; c++) {
        int len = c == 15 ? ny : nz;
        for (int k = 0; k < len; k++) {
            double z = (k - 0.5) * 250.0;
            double v;
            switch (c) {
            case 0: v = z; break;                        /* zu */
            case 1: v = (k - 1.0) * 250.0; break;        /* zw */
            case 2: v = z; break;                        /* zsu */
            case 3: v = (k - 1.0) * 250.0; break;        /* zsw */
            case 4: case 5: v = 1.0; break;              /* mu, mw */
            case 6: case 7: v = 1.1 * exp(-z / 9000.0); break;  /* rhou, rhow */
            case 8: case 9: v = 1.0 / (1.1 * exp(-z / 9000.0)); break;
            case 10: case 11: case 12: v = 300.0 + 0.004 * z; break; /* tb, tbv, tbw */
            case 13: v = moist ? 0.014 * exp(-fmax(z, 0.0) / 2500.0) : 0.0; break; /* qb */
            case 14: v = 1.0 - 9.8 * z / (1004.0 * 300.0); break; /* pib */
            default: v = 1e-4; break;                    /* f */
            }
            col[k] = v;
        }
        orc_set_column(s, c, col);
    }
    const double amp[16] = {5, 5, 1, 1, 1e-3, 1e-3, 1e-3, 50, 10, 2, 0.1, 0.5, 5e-4, 1e-4, 0, 0};
    for (int fld = 0; fld < 14; fld++) {
        if (!moist && (fld == 4 || fld == 5 || fld == 6 || fld == 12)) continue;
        for (size_t i = 0; i < n; i++) {
            double v = amp[fld] * urand();
            if (fld == 5 || fld == 6) v = fabs(v);          /* qc, qr >= 0 */
            if (fld == 13) v = 1.0 - 9.8 * (((int)(i % nz)) - 0.5) * 250.0 / (1004.0 * 300.0) + v; /* pbar */
            f[i] = v;
        }
        for (int role = 0; role < (fld <= 6 ? 3 : 1); role++) orc_set_field(s, fld, role, f);
    }
    orc_step(s, nsteps);
    /* the operators one by one, with the reference's bounds */
    orc_advect_uvw(s, 0.5, 3, nx - 3, 3, ny - 3);
    orc_advect_theta(s, 0.5, 3, nx - 3, 3, ny - 3);
    if (moist) {
        orc_advect_moisture(s, 0.5, 3, nx - 3, 3, ny - 3);
        orc_zero_moisture(s, 1, nx - 1, 1, ny - 1);
        orc_kessler(s, 1, nx - 1, 1, ny - 1);
        orc_apply_bc_micro(s);
        orc_micro_advance_inner(s);
        orc_micro_advance_outer(s);
    }
    orc_solve_pressure(s, 0.5);
    orc_pgf(s, 0.5, 1, nx - 1, 1, ny - 1);
    orc_apply_bc(s);
    orc_rayleigh(s, 1, nx - 1, 1, ny - 1, raylevel);
    orc_advance_inner(s);
    orc_advance_outer(s);
    int bad = 0;
    for (int fld = 0; fld < (moist ? 7 : 4); fld++) {
        orc_get_field(s, fld, 1, f);
        for (size_t i = 0; i < n; i++) bad |= !isfinite(f[i]);
    }
    free(col);
    free(f);
    orc_destroy(s);
    return bad;
}

int main(void) {
    /* transform lengths with every radix of the oracle's FFT, and a prime */
    const int lens[] = {1, 2, 3, 4, 5, 7, 12, 30, 64, 96, 100};
    for (unsigned t = 0; t < sizeof lens / sizeof lens[0]; t++) {
        double *d = malloc(sizeof(double) * 2 * lens[t]);
        for (int i = 0; i < 2 * lens[t]; i++) d[i] = urand();
        orc_fft1d(lens[t], d, -1);
        orc_fft1d(lens[t], d, +1);
        free(d);
    }
    double p[6] = {1, 2, 3, 4, 5, 6};
    (void)orc_interp5(p + 2, 1.0);
    (void)orc_interp3(p + 1, -1.0);
    (void)orc_interp2(p);
    int rc = 0;
    rc |= run(22, 16, 12, 0, 12, 3);   /* dry, damping off */
    rc |= run(22, 16, 12, 1, 8, 3);    /* moist, damping on */
    rc |= run(26, 20, 9, 1, 6, 2);     /* odd sizes */
    printf("oracle sanitizer run: %s\n", rc ? "non-finite state" : "ok");
    return rc;
}
