/*
 * dyncore_oracle.h -- TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement (plain C, FP64) of the serial (PARALLEL 0, periodic east-west)
 * per-timestep dynamical core of michaeldiaz0/Model.  It exists to check the HIP
 * library, never to be shipped or called by it: only tests/, the smoke() entry and
 * bench.py's CPU-baseline leg load it.
 *
 * Field, role and column identifiers match include/dyncore.h so that a test can
 * move arrays between the two without a translation table.
 *
 * What pins each part (the reference ships no tests, golden vectors or fixtures, and its
 * serial build needs fftw3.h, netcdf.h and mpi.h, absent here, so it is not compiled):
 *   orc_fft1d / orc_fft2d      numpy.fft on random data, every transform length the solver
 *                              uses (tests/test_oracle_math.py)
 *   orc_interp5/3/2            exactness for cell averages of polynomials up to the stencil's
 *                              order and the closed form of the upwind term, both signs
 *   orc_poisson_rhs            manufactured solution: the discrete operator applied to a random
 *                              periodic field is inverted to 1e-12, with the (0,0) gauge
 *   orc_kessler_point          per-point conservation of qv+qc+qr, the latent-heat balance,
 *                              saturation after adjustment, inactive-branch identities; each rate
 *                              term against the published Klemp & Wilhelmson (1978) forms in cgs
 *                              units with the SI conversion derived in the test, the saturation
 *                              pressure against IAPWS values, the adjustment as a Newton step
 *                              (tests/test_kessler_published_forms.py)
 *   orc_fall_speed             Klemp & Wilhelmson's 36.34 (rho qr)^0.1364 (rho_0/rho)^1/2 in cgs
 *   orc_advect_* , orc_pgf,    closed-form single-term cases: uniform fields, linear profiles,
 *   orc_solve_pressure,        staggered advecting velocities, buoyancy overwrite, ground flux,
 *   orc_apply_bc, orc_rayleigh projection to zero divergence, sponge/periodic/lid ranges, damping
 *                              profile, rest state (tests/test_operators_closed_form.py)
 *   orc_step                   the composition: a Fourier mode advected by a uniform wind grows
 *                              by the Wicker-Skamarock RK3 polynomial of the published fifth-order
 *                              upwind symbol, over one and three steps, to 1e-12, and stages built
 *                              on the current buffer instead of M fail it
 *                              (tests/test_rk3_fourier_mode.py).  Not compared with the reference's
 *                              own serial build (not buildable here); its written ranges and ghost
 *                              rows follow driver_serial.cpp:178-286 and solver.cpp:229-263.
 */
#ifndef DYNCORE_ORACLE_H
#define DYNCORE_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_grid {
    int nx, ny, nz;          /* NX (6 halo columns included), NY, NZ (ghost levels included) */
    double dx, dz, dt;       /* dy == dx, as the reference parser sets it (process_input.cpp:175-176) */
    int microphysics;        /* 0 = dry, 1 = Kessler warm rain */
    int rayleigh_level;      /* raydampheight; >= nz means off */
    int sponge_north, sponge_south;
} orc_grid;

typedef struct orc_state orc_state;

orc_state *orc_create(const orc_grid *g);
void orc_destroy(orc_state *s);
/* 1 if a fatal condition was hit (Kessler E>100, kessler.cpp:83-88) */
int orc_error(const orc_state *s);

void orc_set_column(orc_state *s, int column, const double *v);
void orc_get_column(const orc_state *s, int column, double *v);
void orc_set_field(orc_state *s, int field, int role, const double *v);
void orc_get_field(const orc_state *s, int field, int role, double *v);

/* whole step: s_integrate_rk3 (driver_serial.cpp:178-286), nsteps times */
void orc_step(orc_state *s, int nsteps);

/* individual reference operators, same bounds convention (half-open, array coordinates) */
void orc_advect_uvw(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_advect_theta(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_advect_moisture(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_zero_moisture(orc_state *s, int il, int ih, int jl, int jh);
void orc_solve_pressure(orc_state *s, double step);
void orc_pgf(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_apply_bc(orc_state *s);
void orc_apply_bc_micro(orc_state *s);
void orc_kessler(orc_state *s, int il, int ih, int jl, int jh);
void orc_rayleigh(orc_state *s, int il, int ih, int jl, int jh, int level);
void orc_advance_inner(orc_state *s);
void orc_advance_outer(orc_state *s);
void orc_micro_advance_inner(orc_state *s);
void orc_micro_advance_outer(orc_state *s);

/* turbulence parameterization, turbulence.cpp: USE_TURBULENT_STRESS on or off, with the land-sea mask
 * landsea[NX*NY] (NULL: all ocean, as initialize_landsea sets it without a file, turbulence.cpp:370-375);
 * when on, orc_step calls calculate_diff_tend and the apply_* functions where s_integrate_rk3 does */
void orc_set_turbulence(orc_state *s, int on, const double *landsea);
void orc_calculate_diff_tend(orc_state *s, int il, int ih, int jl, int jh);
void orc_apply_velocity_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_apply_scalar_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh);
void orc_apply_moisture_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh);
/* 0 Kmix, 1 KHmix, 2-4 u/v/w friction, 5-8 theta/qv/qc/qr diffusion tendencies, [NX][NY][NZ] */
void orc_get_turbulence(const orc_state *s, int which, double *out);
/* Kessler's saturation flags [NX][NY][NZ] (0 or 1) set directly, for tests of the moist branch */
void orc_set_saturated(orc_state *s, const unsigned char *flags);

/* ---- pieces exposed so tests can pin them against mathematics ---- */
/* in-place complex DFT of length n over interleaved (re,im) data; sign -1 forward, +1 backward,
 * unnormalised (the FFTW convention used at pressure.cpp:747-748) */
void orc_fft1d(int n, double *data, int sign);
/* 2-D complex DFT of `howmany` planes of n0 x n1 (row-major, n1 fastest), unnormalised */
void orc_fft2d(int n0, int n1, int howmany, double *data, int sign);
/* face interpolants of fluxes.cpp:110-206 on a 1-D array p with unit stride, face between p[0] and p[1] */
double orc_interp5(const double *p, double sign);
double orc_interp3(const double *p, double sign);
double orc_interp2(const double *p);
/* Poisson solve on a given real right-hand side rhs[k][i][j] (k in 0..nz-1, i in 0..ni-1, j in 0..ny-1;
 * only k=1..nz-2 are used), result x[k][i][j] at k=1..nz-2; ni = nx-6 */
void orc_poisson_rhs(orc_state *s, const double *rhs, double *x);
/* one Kessler point: in/out th, qv, qc, qr (perturbations), returns E>100 flag */
int orc_kessler_point(double dt, double rhou, double tb, double qb, double tbv,
                      double thbar, double qbar, double pbar, double pi,
                      double *th, double *qv, double *qc, double *qr);
/* rain fall speed at one point (kessler.cpp:222-252) */
double orc_fall_speed(double qr, double rhou_k, double rhou_1);

#ifdef __cplusplus
}
#endif
#endif
