/*
 * dyncore_oracle.c -- TEST INFRASTRUCTURE ONLY (the oracle).
 *
 * A plain-C, FP64 restatement of the serial per-timestep dynamical core of
 * michaeldiaz0/Model (PARALLEL 0, periodic east-west boundaries, non-hydrostatic,
 * non-linear, HOR_ADVECTION_ORDER 5, VER_ADVECTION_ORDER 3, STRETCHED_GRID 1), used
 * only to check the HIP library in tests/, smoke() and bench.py's CPU-baseline leg.
 *
 * Every function follows the reference file:line it cites.  Expression trees are kept
 * in the reference's evaluation order; build with -ffp-contract=off for parity.
 *
 * Restatement choices (mathematically identical to the reference):
 *  - The reference builds face fluxes slab by slab into struct cell/vel_cell scratch and
 *    carries the east face of slab i-1 over as the west face of slab i
 *    (fluxes.cpp:891,1322,1504).  Every face value is a pure function of its position, so
 *    here a face is recomputed where it is needed; the values are bit-identical.
 *  - Sign cells (compute_sign_cells, fluxes.cpp:316-341) are recomputed from U,V,W on use;
 *    nothing between compute_sign_cells and their last use writes U, V or W.
 *  - The 2-D DFT is applied only to levels 1..NZ-2: levels 0 and NZ-1 of out3d are never read
 *    by the tridiagonal solve (pressure.cpp:837-871), and PI at 0 / NZ-1 comes from levels 1 /
 *    NZ-2 (pressure.cpp:881,890).
 *  - The FFT is this file's own mixed-radix Cooley-Tukey (radix 2..5 + generic DFT factor),
 *    not FFTW; results agree with any exact DFT to rounding.
 *
 * Parity status of each piece is recorded in the tests that pin it.
 */
#include "dyncore_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* identifiers shared with include/dyncore.h */
enum { F_U = 0, F_V, F_W, F_TH, F_QV, F_QC, F_QR, F_PI, F_UBAR, F_VBAR, F_WBAR, F_THBAR,
       F_QBAR, F_PBAR, F_VT, F_ACCRAIN, F_COUNT };
enum { R_M = 0, R_CUR = 1, R_P = 2 };
enum { C_ZU = 0, C_ZW, C_ZSU, C_ZSW, C_MU, C_MW, C_RHOU, C_RHOW, C_ONE_D_RHOU, C_ONE_D_RHOW,
       C_TB, C_TBV, C_TBW, C_QB, C_PIB, C_F, C_COUNT };

/* physical constants, model.h:250-258 */
static const double Rd = 287.0;
static const double cp = 1004;
static const double p0 = 100000;
static const double grav = 9.8;
static const double Lv = 2.5e6;
static const double trigpi = 3.14159265358979324;

/* interpolation weights, fluxes.cpp:218-222 */
static const double onetwelfth = 1.0 / 12.0;
static const double seventwelfth = 7.0 / 12.0;
static const double thirtyseven60ths = 37.0 / 60.0;
static const double two15ths = 2.0 / 15.0;
static const double one60ths = 1.0 / 60.0;

static const double minvar = 1.0e-12; /* kessler.cpp:6 */

typedef struct { double re, im; } cplx;

struct orc_state {
    int NX, NY, NZ, NYNZ;
    double dx, dy, dz, dt, dtx, dty, dtz, one_d_dx, one_d_dy, one_d_dz;
    int moist, raylevel, jn, js;
    int error;
    double *col[C_COUNT];
    /* three rotating buffers per prognostic (solver.cpp:229-263, microphysics.cpp:642-669) */
    double *us, *ups, *ums, *vs, *vps, *vms, *ws, *wps, *wms, *ths, *thps, *thms;
    double *qvs, *qvps, *qvms, *qcs, *qcps, *qcms, *qrs, *qrps, *qrms;
    double *pis, *vts, *accrain;
    double *ubar, *vbar, *wbar, *thbar, *qbar, *pbar;
    /* pressure solver (init_fft3d, pressure.cpp:724-797) */
    int ni, nj;
    double *kwave, *lwave, *acoef, *ccoef, *bcoef, *rhs;
    cplx *in3d;
    /* sponge coefficients (init_boundaries, boundaries.cpp:44-114) */
    double *north_c, *south_c;
    /* turbulence parameterization (turbulence.cpp:32-60, 85-147): mixing coefficients, stored
     * tendencies, the land-sea mask and the saturation flags Kessler leaves (kessler.cpp:129-150) */
    int turb;
    double *kmix, *khmix, *ufric, *vfric, *wfric, *tdiff, *qvdiff, *qcdiff, *qrdiff, *landsea;
    unsigned char *saturated;
    double *vml2, *vkmax;
    double kmixh_max, kmixv_max, lH;
};

#define IDX(i, j, k) ((size_t)(i) * (size_t)s->NYNZ + (size_t)(j) * (size_t)s->NZ + (size_t)(k))

/* ------------------------------------------------------------------------------------------------
 * FFT (project-written; stands in for FFTW's fftw_plan_many_dft, pressure.cpp:747-748)
 * ----------------------------------------------------------------------------------------------*/
static int smallest_factor(int n) {
    static const int pref[] = {4, 2, 3, 5};
    for (int t = 0; t < 4; t++)
        if (n % pref[t] == 0) return pref[t];
    for (int p = 7; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* out[m] = sum_n in[n*istride] * exp(sign*2*pi*i*n*m/N); tw holds exp(sign*2*pi*i*e/Ntot) */
static void fft_rec(int N, const cplx *in, int istride, cplx *out, const cplx *tw, int Ntot, cplx *scratch) {
    if (N == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(N);
    int M = N / p;
    int twstride = Ntot / N;
    for (int r = 0; r < p; r++) fft_rec(M, in + (size_t)r * istride, istride * p, out + (size_t)r * M, tw, Ntot, scratch);
    /* combine: X[m + q*M] = sum_r W_N^{r(m+qM)} * sub_r[m] */
    for (int m = 0; m < M; m++) {
        for (int q = 0; q < p; q++) {
            int mm = m + q * M;
            double re = 0, im = 0;
            for (int r = 0; r < p; r++) {
                long e = ((long)r * mm) % N;
                cplx w = tw[e * twstride];
                cplx x = out[r * M + m];
                re += w.re * x.re - w.im * x.im;
                im += w.re * x.im + w.im * x.re;
            }
            scratch[mm].re = re;
            scratch[mm].im = im;
        }
    }
    memcpy(out, scratch, sizeof(cplx) * (size_t)N);
}

static cplx *make_twiddles(int N, int sign) {
    cplx *tw = (cplx *)malloc(sizeof(cplx) * (size_t)N);
    for (int e = 0; e < N; e++) {
        double a = 2.0 * trigpi * (double)e / (double)N;
        tw[e].re = cos(a);
        tw[e].im = sign * sin(a);
    }
    return tw;
}

/* strided 1-D transform of `count` lines using precomputed twiddles */
static void fft_lines(int N, cplx *base, size_t stride, size_t line_step, int count, const cplx *tw) {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)N * 3);
    cplx *in = buf, *out = buf + N, *scr = buf + 2 * N;
    for (int l = 0; l < count; l++) {
        cplx *p = base + (size_t)l * line_step;
        for (int n = 0; n < N; n++) in[n] = p[(size_t)n * stride];
        fft_rec(N, in, 1, out, tw, N, scr);
        for (int n = 0; n < N; n++) p[(size_t)n * stride] = out[n];
    }
    free(buf);
}

void orc_fft1d(int n, double *data, int sign) {
    cplx *tw = make_twiddles(n, sign);
    fft_lines(n, (cplx *)data, 1, 0, 1, tw);
    free(tw);
}

/* planes of n0 x n1, n1 fastest: transform along n1 then along n0 */
static void fft2d_planes(int n0, int n1, int howmany, cplx *data, const cplx *tw0, const cplx *tw1) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int h = 0; h < howmany; h++) {
        cplx *plane = data + (size_t)h * n0 * n1;
        fft_lines(n1, plane, 1, (size_t)n1, n0, tw1);
        fft_lines(n0, plane, (size_t)n1, 1, n1, tw0);
    }
}

void orc_fft2d(int n0, int n1, int howmany, double *data, int sign) {
    cplx *tw0 = make_twiddles(n0, sign), *tw1 = make_twiddles(n1, sign);
    fft2d_planes(n0, n1, howmany, (cplx *)data, tw0, tw1);
    free(tw0);
    free(tw1);
}

/* ------------------------------------------------------------------------------------------------
 * Face interpolants, fluxes.cpp:110-206.  p points at the cell on the low side of the face,
 * st is the element stride along the interpolation direction.
 * ----------------------------------------------------------------------------------------------*/
static inline double interp6(const double *p, ptrdiff_t st) {
    /* i_interp6th(s,i+1,i,i+2,i-1,i+3,i-2), fluxes.cpp:110,155-157 */
    return thirtyseven60ths * (p[st] + p[0]) - two15ths * (p[2 * st] + p[-st]) + one60ths * (p[3 * st] + p[-2 * st]);
}
static inline double interp5(const double *p, ptrdiff_t st, double sign) {
    /* INTERP_5TH_EAST, fluxes.cpp:117,170-172 */
    return interp6(p, st) + one60ths * sign * (-(p[3 * st] - p[-2 * st]) + 5.0 * (p[2 * st] - p[-st]) - 10.0 * (p[st] - p[0]));
}
static inline double interp3(const double *p, ptrdiff_t st, double sign) {
    /* INTERP_3RD_TOP, fluxes.cpp:135,191-199 */
    return (seventwelfth * (p[st] + p[0]) - onetwelfth * (p[2 * st] + p[-st])) +
           onetwelfth * sign * ((p[2 * st] - p[-st]) - 3.0 * (p[st] - p[0]));
}
static inline double interp2(const double *p, ptrdiff_t st) { return 0.5 * (p[st] + p[0]); } /* fluxes.cpp:206 */

double orc_interp5(const double *p, double sign) { return interp5(p, 1, sign); }
double orc_interp3(const double *p, double sign) { return interp3(p, 1, sign); }
double orc_interp2(const double *p) { return interp2(p, 1); }

/* signof, fluxes.cpp:295-311: zero maps to -1 */
static inline double signof(double x) { return x > 0 ? 1.0 : -1.0; }

/* ------------------------------------------------------------------------------------------------
 * Momentum control-volume faces (interpolate_velocity + compute_fluxes, fluxes.cpp:582-627,1109-1592)
 * ----------------------------------------------------------------------------------------------*/
typedef struct { double e, n, t; } faces;

/* u control volume at (i,j,k): east face (i+1/2), north face (j+1/2), top face (k+1/2) fluxes */
static faces u_faces(const orc_state *s, int i, int j, int k) {
    const double *U = s->us, *V = s->vs, *W = s->ws, *UB = s->ubar, *VB = s->vbar, *WB = s->wbar;
    const ptrdiff_t si = s->NYNZ, sj = s->NZ;
    faces f;
    /* advecting velocities, fluxes.cpp:1430-1476 */
    double ua = 0.5 * (U[IDX(i + 1, j, k)] + U[IDX(i, j, k)]);
    double va = 0.5 * (V[IDX(i, j + 1, k)] + V[IDX(i - 1, j + 1, k)]);
    double uba = 0.5 * (UB[IDX(i + 1, j, k)] + UB[IDX(i, j, k)]);
    double vba = 0.5 * (VB[IDX(i, j + 1, k)] + VB[IDX(i - 1, j + 1, k)]);
    /* horizontal interpolants and fluxes, fluxes.cpp:1492-1511, 595-597 */
    double ue = interp5(U + IDX(i, j, k), si, signof(uba + ua));
    double ube = interp5(UB + IDX(i, j, k), si, signof(ua));
    double un = interp5(U + IDX(i, j, k), sj, signof(vba + va));
    double ubn = interp5(UB + IDX(i, j, k), sj, signof(va));
    f.e = ua * ube + (uba + ua) * ue;
    f.n = va * ubn + (vba + va) * un;
    if (k == 0) { f.t = 0; return f; } /* fluxes.cpp:588 */
    double wa = 0.5 * (W[IDX(i, j, k + 1)] + W[IDX(i - 1, j, k + 1)]);
    double wba = 0.5 * (WB[IDX(i, j, k + 1)] + WB[IDX(i - 1, j, k + 1)]);
    double ut, ubt;
    if (k == 1 || k == s->NZ - 2) { /* interpolate_velocity_vertical_2nd, fluxes.cpp:1170-1188 */
        ut = interp2(U + IDX(i, j, k), 1);
        ubt = interp2(UB + IDX(i, j, k), 1);
    } else { /* INTERP_3RD_TOP, fluxes.cpp:1507,1511 */
        ut = interp3(U + IDX(i, j, k), 1, signof(wba + wa));
        ubt = interp3(UB + IDX(i, j, k), 1, signof(wa));
    }
    f.t = wa * ubt + (wba + wa) * ut;
    return f;
}

/* v control volume at (i,j,k) */
static faces v_faces(const orc_state *s, int i, int j, int k) {
    const double *U = s->us, *V = s->vs, *W = s->ws, *UB = s->ubar, *VB = s->vbar, *WB = s->wbar;
    const ptrdiff_t si = s->NYNZ, sj = s->NZ;
    faces f;
    double ua = 0.5 * (U[IDX(i + 1, j, k)] + U[IDX(i + 1, j - 1, k)]);
    double va = 0.5 * (V[IDX(i, j + 1, k)] + V[IDX(i, j, k)]);
    double uba = 0.5 * (UB[IDX(i + 1, j, k)] + UB[IDX(i + 1, j - 1, k)]);
    double vba = 0.5 * (VB[IDX(i, j + 1, k)] + VB[IDX(i, j, k)]);
    double ve = interp5(V + IDX(i, j, k), si, signof(uba + ua));
    double vbe = interp5(VB + IDX(i, j, k), si, signof(ua));
    double vn = interp5(V + IDX(i, j, k), sj, signof(vba + va));
    double vbn = interp5(VB + IDX(i, j, k), sj, signof(va));
    f.e = ua * vbe + (uba + ua) * ve;
    f.n = va * vbn + (vba + va) * vn;
    if (k == 0) { f.t = 0; return f; }
    double wa = 0.5 * (W[IDX(i, j, k + 1)] + W[IDX(i, j - 1, k + 1)]);
    double wba = 0.5 * (WB[IDX(i, j, k + 1)] + WB[IDX(i, j - 1, k + 1)]);
    double vt, vbt;
    if (k == 1 || k == s->NZ - 2) {
        vt = interp2(V + IDX(i, j, k), 1);
        vbt = interp2(VB + IDX(i, j, k), 1);
    } else {
        vt = interp3(V + IDX(i, j, k), 1, signof(wba + wa));
        vbt = interp3(VB + IDX(i, j, k), 1, signof(wa));
    }
    f.t = wa * vbt + (wba + wa) * vt;
    return f;
}

/* ------------------------------------------------------------------------------------------------
 * Slab loop shared by the advection operators: the reference marches i-slab by i-slab, computes
 * every face of slab i once (compute_fluxes*, rows jl-1..jh-1) and reuses slab i-1's east face as
 * the west face (compute_west* for the first slab).  Chunks of i run on separate OpenMP threads in
 * the timed build; each chunk starts with its own compute_west, so results do not depend on the
 * thread count.
 * ----------------------------------------------------------------------------------------------*/
typedef void (*slab_faces_fn)(const orc_state *s, int i, int jl, int jh, double *const *face);
typedef void (*slab_update_fn)(orc_state *s, double step, int i, int jl, int jh, double *const *face, double *const *west);

static void slab_march(orc_state *s, double step, int il, int ih, int jl, int jh, int nface, slab_faces_fn faces_fn,
                       slab_update_fn update_fn) {
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
        int tid = 0, nth = 1;
#ifdef _OPENMP
        tid = omp_get_thread_num();
        nth = omp_get_num_threads();
#endif
        int n = ih - il, i0 = il + (int)((long)n * tid / nth), i1 = il + (int)((long)n * (tid + 1) / nth);
        if (i1 > i0) {
            size_t plane = (size_t)s->NYNZ;
            double *mem = (double *)calloc(plane * (size_t)nface * 2, sizeof(double));
            double *face[8], *west[8];
            for (int f = 0; f < nface; f++) { face[f] = mem + plane * f; west[f] = mem + plane * (nface + f); }
            faces_fn(s, i0 - 1, jl - 1, jh, west); /* compute_west* */
            for (int i = i0; i < i1; i++) {
                faces_fn(s, i, jl - 1, jh, face);  /* compute_fluxes* */
                update_fn(s, step, i, jl, jh, face, west);
                for (int f = 0; f < nface; f++) { double *t = west[f]; west[f] = face[f]; face[f] = t; }
            }
            free(mem);
        }
    }
}
#define PL(j, k) ((size_t)(j) * (size_t)s->NZ + (size_t)(k))

/* ------------------------------------------------------------------------------------------------
 * advect_uvw_velocity, advection.cpp:304-339 (u_tend_cell 73-94, v_tend_cell 101-122,
 * buoyancy 42-51, 329-334).  The w flux tendency is overwritten by the buoyancy update
 * (advection.cpp:323 then 329-334), so it is not formed.
 * face[] = {u east, u north, u top, v east, v north, v top}
 * ----------------------------------------------------------------------------------------------*/
static void uv_slab_faces(const orc_state *s, int i, int jl, int jh, double *const *fc) {
    for (int j = jl; j < jh; j++) {
        fc[2][PL(j, 0)] = 0; /* fluxes.cpp:588-589 */
        fc[5][PL(j, 0)] = 0;
        for (int k = 1; k < s->NZ - 1; k++) {
            faces u = u_faces(s, i, j, k), v = v_faces(s, i, j, k);
            fc[0][PL(j, k)] = u.e; fc[1][PL(j, k)] = u.n; fc[2][PL(j, k)] = u.t;
            fc[3][PL(j, k)] = v.e; fc[4][PL(j, k)] = v.n; fc[5][PL(j, k)] = v.t;
        }
    }
}

static void uv_slab_update(orc_state *s, double step, int i, int jl, int jh, double *const *fc, double *const *wst) {
    const double *rhow = s->col[C_RHOW], *one_d_rhou = s->col[C_ONE_D_RHOU], *mu = s->col[C_MU], *f = s->col[C_F];
    const double *U = s->us, *V = s->vs;
    for (int j = jl; j < jh; j++)
        for (int k = 1; k < s->NZ - 1; k++) {
            double DTZ = s->dtz * mu[k];
            double uother = 0.25 * (V[IDX(i, j, k)] + V[IDX(i, j + 1, k)] + V[IDX(i - 1, j, k)] + V[IDX(i - 1, j + 1, k)]);
            double tu = -s->dtx * (fc[0][PL(j, k)] - wst[0][PL(j, k)]) - s->dty * (fc[1][PL(j, k)] - fc[1][PL(j - 1, k)]) -
                        DTZ * (rhow[k + 1] * fc[2][PL(j, k)] - rhow[k] * fc[2][PL(j, k - 1)]) * one_d_rhou[k] + s->dt * f[j] * uother;
            s->ups[IDX(i, j, k)] = s->ums[IDX(i, j, k)] + step * tu;
        }
    for (int j = jl; j < jh; j++)
        for (int k = 1; k < s->NZ - 1; k++) {
            double DTZ = s->dtz * mu[k];
            double vother = 0.25 * (U[IDX(i, j, k)] + U[IDX(i + 1, j, k)] + U[IDX(i, j - 1, k)] + U[IDX(i + 1, j - 1, k)]);
            double tv = -s->dtx * (fc[3][PL(j, k)] - wst[3][PL(j, k)]) - s->dty * (fc[4][PL(j, k)] - fc[4][PL(j - 1, k)]) -
                        DTZ * (rhow[k + 1] * fc[5][PL(j, k)] - rhow[k] * fc[5][PL(j, k - 1)]) * one_d_rhou[k] - s->dt * f[j] * vother;
            s->vps[IDX(i, j, k)] = s->vms[IDX(i, j, k)] + step * tv;
        }
}

void orc_advect_uvw(orc_state *s, double step, int il, int ih, int jl, int jh) {
    slab_march(s, step, il, ih, jl, jh, 6, uv_slab_faces, uv_slab_update);
    const double *tbv = s->col[C_TBV];
    /* buoyancy (advection.cpp:329-334), after the slab loop */
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int k = 2; k < s->NZ - 1; k++) {
                double bk, bkm;
                if (s->moist == 0) {
                    bk = s->ths[IDX(i, j, k)] / tbv[k];
                    bkm = s->ths[IDX(i, j, k - 1)] / tbv[k - 1];
                } else {
                    bk = s->ths[IDX(i, j, k)] / tbv[k] + 0.61 * s->qvs[IDX(i, j, k)] - s->qcs[IDX(i, j, k)] - s->qrs[IDX(i, j, k)];
                    bkm = s->ths[IDX(i, j, k - 1)] / tbv[k - 1] + 0.61 * s->qvs[IDX(i, j, k - 1)] - s->qcs[IDX(i, j, k - 1)] -
                          s->qrs[IDX(i, j, k - 1)];
                }
                s->wps[IDX(i, j, k)] = s->wms[IDX(i, j, k)] + step * s->dt * grav * 0.5 * (bk + bkm);
            }
}

/* ------------------------------------------------------------------------------------------------
 * Scalar with basic state (theta, qv): interpolate_scalar (797-867, 878-942) + compute_fluxes_scalar
 * (632-654).  Perturbation interpolants take the sign of the total wind (SIGN_B_VEL), basic-state
 * interpolants the sign of the perturbation wind (SIGN_P_VEL).
 * ----------------------------------------------------------------------------------------------*/
static double sc_east(const orc_state *s, const double *S, const double *SB, int i, int j, int k) {
    double U1 = s->us[IDX(i + 1, j, k)], UB1 = s->ubar[IDX(i + 1, j, k)];
    double se = interp5(S + IDX(i, j, k), s->NYNZ, signof(UB1 + U1));
    double be = interp5(SB + IDX(i, j, k), s->NYNZ, signof(U1));
    return U1 * be + (UB1 + U1) * se;
}
static double sc_north(const orc_state *s, const double *S, const double *SB, int i, int j, int k) {
    double V1 = s->vs[IDX(i, j + 1, k)], VB1 = s->vbar[IDX(i, j + 1, k)];
    double sn = interp5(S + IDX(i, j, k), s->NZ, signof(VB1 + V1));
    double bn = interp5(SB + IDX(i, j, k), s->NZ, signof(V1));
    return V1 * bn + (VB1 + V1) * sn;
}
static double sc_top(const orc_state *s, const double *S, const double *SB, int i, int j, int k) {
    if (k == 0) return 0.0; /* fluxes.cpp:638 */
    double W1 = s->ws[IDX(i, j, k + 1)], WB1 = s->wbar[IDX(i, j, k + 1)];
    double st, bt;
    if (k == 1 || k == s->NZ - 2) { /* fluxes.cpp:829-836, 903-912 */
        st = interp2(S + IDX(i, j, k), 1);
        bt = interp2(SB + IDX(i, j, k), 1);
    } else { /* fluxes.cpp:861-865, 936-940 */
        st = interp3(S + IDX(i, j, k), 1, signof(WB1 + W1));
        bt = interp3(SB + IDX(i, j, k), 1, signof(W1));
    }
    return W1 * bt + (WB1 + W1) * st;
}

/* Z_ADVECT_BASE, advection.cpp:16-18 */
static inline double z_advect_base(const orc_state *s, const double *base, int i, int j, int k) {
    const double *rhow = s->col[C_RHOW];
    return (0.50 * (rhow[k + 1] * s->ws[IDX(i, j, k + 1)] * (base[k + 1] - base[k]) + rhow[k] * s->ws[IDX(i, j, k)] * (base[k] - base[k - 1])) *
            s->col[C_ONE_D_RHOU][k]);
}

/* advect_theta, advection.cpp:349-373 with theta_tend_cell 152-168; face[] = {east, north, top} */
static void th_slab_faces(const orc_state *s, int i, int jl, int jh, double *const *fc) {
    for (int j = jl; j < jh; j++) {
        fc[2][PL(j, 0)] = 0;
        for (int k = 1; k < s->NZ - 1; k++) {
            fc[0][PL(j, k)] = sc_east(s, s->ths, s->thbar, i, j, k);
            fc[1][PL(j, k)] = sc_north(s, s->ths, s->thbar, i, j, k);
            fc[2][PL(j, k)] = sc_top(s, s->ths, s->thbar, i, j, k);
        }
    }
}
static void th_slab_update(orc_state *s, double step, int i, int jl, int jh, double *const *fc, double *const *wst) {
    const double *rhow = s->col[C_RHOW], *one_d_rhou = s->col[C_ONE_D_RHOU], *mu = s->col[C_MU], *tb = s->col[C_TB];
    for (int j = jl; j < jh; j++)
        for (int k = 1; k < s->NZ - 1; k++) {
            double DTZ = s->dtz * mu[k];
            double t = -s->dtx * (fc[0][PL(j, k)] - wst[0][PL(j, k)]) - s->dty * (fc[1][PL(j, k)] - fc[1][PL(j - 1, k)]) -
                       DTZ * (rhow[k + 1] * fc[2][PL(j, k)] - rhow[k] * fc[2][PL(j, k - 1)]) * one_d_rhou[k] - DTZ * z_advect_base(s, tb, i, j, k);
            s->thps[IDX(i, j, k)] = s->thms[IDX(i, j, k)] + step * t;
        }
}
void orc_advect_theta(orc_state *s, double step, int il, int ih, int jl, int jh) {
    slab_march(s, step, il, ih, jl, jh, 3, th_slab_faces, th_slab_update);
}

/* ------------------------------------------------------------------------------------------------
 * Moisture: advect_microphysics_cell, advection.cpp:451-498; compute_fluxes_moisture,
 * fluxes.cpp:719-784; interpolate_scalar (878-942); interpolate_scalar_with_fallspeed (953-1032);
 * fall speed calculate_rain_fall_speed_kessler, kessler.cpp:222-252.
 * ----------------------------------------------------------------------------------------------*/
double orc_fall_speed(double qr, double rhou_k, double rhou_1) {
    double vt = 0.0;
    if (qr > minvar) {
        double qrr = fmax(qr * 0.001 * rhou_k, 0.0);
        double vtden = sqrt(rhou_1 / rhou_k);
        vt = 36.34 * (pow(qrr, 0.1364)) * vtden;
    }
    return vt;
}

static void fall_speed_field(orc_state *s, int il, int ih, int jl, int jh) {
    const double *rhou = s->col[C_RHOU];
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++) {
            for (int k = 1; k < s->NZ - 1; k++) s->vts[IDX(i, j, k)] = orc_fall_speed(s->qrs[IDX(i, j, k)], rhou[k], rhou[1]);
            s->vts[IDX(i, j, 0)] = s->vts[IDX(i, j, 1)]; /* kessler.cpp:247-248 (HTOPO = 0) */
        }
}

/* perturbation-only scalar faces (qc, qr horizontal; qc vertical) */
static double q_east(const orc_state *s, const double *S, int i, int j, int k) {
    double ub = s->ubar[IDX(i + 1, j, k)] + s->us[IDX(i + 1, j, k)];
    return ub * interp5(S + IDX(i, j, k), s->NYNZ, signof(ub));
}
static double q_north(const orc_state *s, const double *S, int i, int j, int k) {
    double vb = s->vbar[IDX(i, j + 1, k)] + s->vs[IDX(i, j + 1, k)];
    return vb * interp5(S + IDX(i, j, k), s->NZ, signof(vb));
}
static double qc_top(const orc_state *s, int i, int j, int k) {
    if (k == 0) return 0.0; /* fluxes.cpp:733 */
    double wb = s->wbar[IDX(i, j, k + 1)] + s->ws[IDX(i, j, k + 1)];
    double t = (k == 1 || k == s->NZ - 2) ? interp2(s->qcs + IDX(i, j, k), 1) : interp3(s->qcs + IDX(i, j, k), 1, signof(wb));
    return wb * t;
}
static double qr_top(const orc_state *s, int i, int j, int k) {
    const double *VT = s->vts, *QR = s->qrs;
    if (k == 0) { /* fluxes.cpp:966, 734: rain leaves through the ground */
        double t0 = 0.5 * (QR[IDX(i, j, 0)] + QR[IDX(i, j, 1)]);
        return -0.5 * (VT[IDX(i, j, 1)] + VT[IDX(i, j, 0)]) * t0;
    }
    double wb = s->wbar[IDX(i, j, k + 1)] + s->ws[IDX(i, j, k + 1)];
    double t;
    if (k == 1 || k == s->NZ - 2) t = interp2(QR + IDX(i, j, k), 1);
    else t = interp3(QR + IDX(i, j, k), 1, signof(s->wbar[IDX(i, j, k + 1)] + s->ws[IDX(i, j, k + 1)] - 0.5 * (VT[IDX(i, j, k)] + VT[IDX(i, j, k + 1)])));
    return (wb - 0.5 * (VT[IDX(i, j, k + 1)] + VT[IDX(i, j, k)])) * t;
}

/* face[] = {qv east, qv north, qv top, qc east, qc north, qc top, qr east, qr north} + qr top below */
static void q_slab_faces(const orc_state *s, int i, int jl, int jh, double *const *fc) {
    for (int j = jl; j < jh; j++) {
        fc[2][PL(j, 0)] = 0;
        fc[5][PL(j, 0)] = 0;
        for (int k = 1; k < s->NZ - 1; k++) {
            fc[0][PL(j, k)] = sc_east(s, s->qvs, s->qbar, i, j, k);
            fc[1][PL(j, k)] = sc_north(s, s->qvs, s->qbar, i, j, k);
            fc[2][PL(j, k)] = sc_top(s, s->qvs, s->qbar, i, j, k);
            fc[3][PL(j, k)] = q_east(s, s->qcs, i, j, k);
            fc[4][PL(j, k)] = q_north(s, s->qcs, i, j, k);
            fc[5][PL(j, k)] = qc_top(s, i, j, k);
            fc[6][PL(j, k)] = q_east(s, s->qrs, i, j, k);
            fc[7][PL(j, k)] = q_north(s, s->qrs, i, j, k);
        }
    }
}
static void q_slab_update(orc_state *s, double step, int i, int jl, int jh, double *const *fc, double *const *wst) {
    const double *rhow = s->col[C_RHOW], *one_d_rhou = s->col[C_ONE_D_RHOU], *mu = s->col[C_MU], *qb = s->col[C_QB];
    for (int j = jl; j < jh; j++) {
        double qrtop_below = qr_top(s, i, j, 0);
        for (int k = 1; k < s->NZ - 1; k++) {
            double DTZ = s->dtz * mu[k];
            /* vapour: qv_tend_cell, advection.cpp:197-213 */
            double t = -s->dtx * (fc[0][PL(j, k)] - wst[0][PL(j, k)]) - s->dty * (fc[1][PL(j, k)] - fc[1][PL(j - 1, k)]) -
                       DTZ * (rhow[k + 1] * fc[2][PL(j, k)] - rhow[k] * fc[2][PL(j, k - 1)]) * one_d_rhou[k] - DTZ * z_advect_base(s, qb, i, j, k);
            s->qvps[IDX(i, j, k)] = s->qvms[IDX(i, j, k)] + step * t;
            /* cloud: qc_tend_cell, advection.cpp:220-233 */
            t = -s->dtx * (fc[3][PL(j, k)] - wst[3][PL(j, k)]) - s->dty * (fc[4][PL(j, k)] - fc[4][PL(j - 1, k)]) -
                DTZ * (rhow[k + 1] * fc[5][PL(j, k)] - rhow[k] * fc[5][PL(j, k - 1)]) * one_d_rhou[k];
            s->qcps[IDX(i, j, k)] = s->qcms[IDX(i, j, k)] + step * t;
            /* rain: qr_tend_cell, advection.cpp:240-253 */
            double qrtop = qr_top(s, i, j, k);
            t = -s->dtx * (fc[6][PL(j, k)] - wst[6][PL(j, k)]) - s->dty * (fc[7][PL(j, k)] - fc[7][PL(j - 1, k)]) -
                DTZ * (rhow[k + 1] * qrtop - rhow[k] * qrtop_below) * one_d_rhou[k];
            s->qrps[IDX(i, j, k)] = s->qrms[IDX(i, j, k)] + step * t;
            qrtop_below = qrtop;
        }
    }
}

void orc_advect_moisture(orc_state *s, double step, int il, int ih, int jl, int jh) {
    fall_speed_field(s, il, ih, jl, jh); /* advection.cpp:456-457 */
    slab_march(s, step, il, ih, jl, jh, 8, q_slab_faces, q_slab_update);
}

/* zero_moisture, microphysics.cpp:675-700 */
void orc_zero_moisture(orc_state *s, int il, int ih, int jl, int jh) {
    size_t size = (size_t)s->NX * s->NYNZ;
    for (size_t n = 0; n < size; n++)
        if (s->qcps[n] < 0) s->qcps[n] = 0;
    for (size_t n = 0; n < size; n++)
        if (s->qrps[n] < 0) s->qrps[n] = 0;
    const double *qb = s->col[C_QB];
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int k = 1; k < s->NZ - 1; k++)
                if (s->qvps[IDX(i, j, k)] + s->qbar[IDX(i, j, k)] + qb[k] < 0) s->qvps[IDX(i, j, k)] = -(s->qbar[IDX(i, j, k)] + qb[k]);
}

/* ------------------------------------------------------------------------------------------------
 * Pressure: init_fft3d (pressure.cpp:724-797), poisson_fft3d (805-893), divergence3d_periodic
 * (1390-1424), periodic boundaries (boundaries.cpp:533-555).
 * ----------------------------------------------------------------------------------------------*/
#define IN3D(i, j, k) s->in3d[(size_t)nj * ((size_t)(i) + (size_t)ni * (size_t)(k)) + (size_t)(j)]
#define B2(i, j, k) s->bcoef[(size_t)nk * ((size_t)(j) + (size_t)nj * (size_t)(i)) + (size_t)(k)]
#define F2(i, j, k) s->rhs[(size_t)nk * ((size_t)(j) + (size_t)nj * (size_t)(i)) + (size_t)(k)]

static void init_pressure_solver(orc_state *s) {
    int ni = s->NX - 6, nj = s->NY, nk = s->NZ;
    const double dx = s->dx, dz = s->dz;
    s->ni = ni;
    s->nj = nj;
    s->kwave = (double *)malloc(sizeof(double) * ni);
    s->lwave = (double *)malloc(sizeof(double) * nj);
    s->acoef = (double *)calloc(nk, sizeof(double));
    s->ccoef = (double *)calloc(nk, sizeof(double));
    s->bcoef = (double *)calloc((size_t)ni * nj * nk, sizeof(double));
    s->rhs = (double *)calloc((size_t)ni * nj * nk, sizeof(double));
    s->in3d = (cplx *)calloc((size_t)ni * nj * nk, sizeof(cplx));
}

/* coefficient set-up, pressure.cpp:751-752, 779-792 (STRETCHED_GRID branch) */
static void pressure_coefficients(orc_state *s) {
    int ni = s->ni, nj = s->nj, nk = s->NZ;
    const double dx = s->dx, dz = s->dz;
    const double *mu = s->col[C_MU], *mw = s->col[C_MW], *rhow = s->col[C_RHOW], *rhou = s->col[C_RHOU];
    for (int i = 0; i < ni; i++) s->kwave[i] = (double)i * 2 * trigpi / ((double)ni);
    for (int i = 0; i < nj; i++) s->lwave[i] = (double)i * 2 * trigpi / ((double)nj);
    for (int k = 1; k < nk - 1; k++) {
        s->acoef[k] = mu[k] * mw[k] * rhow[k] / (dz * dz);
        s->ccoef[k] = mu[k] * mw[k + 1] * rhow[k + 1] / (dz * dz);
    }
    s->acoef[1] = 0;
    s->ccoef[nk - 2] = 0;
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            B2(i, j, 1) = rhou[1] * 2.0 * (cos(s->kwave[i]) + cos(s->lwave[j]) - 2.0) / (dx * dx) - mu[1] * mw[2] * rhow[2] / (dz * dz);
            for (int k = 2; k < nk - 2; k++)
                B2(i, j, k) = rhou[k] * 2.0 * (cos(s->kwave[i]) + cos(s->lwave[j]) - 2.0) / (dx * dx) -
                              mu[k] * (mw[k + 1] * rhow[k + 1] + mw[k] * rhow[k]) / (dz * dz);
            B2(i, j, nk - 2) = rhou[nk - 2] * 2.0 * (cos(s->kwave[i]) + cos(s->lwave[j]) - 2.0) / (dx * dx) -
                               (mu[nk - 2] * mw[nk - 2] * rhow[nk - 2]) / (dz * dz);
        }
}

#define PERIODIC_EW(A)                                        \
    do {                                                      \
        A[IDX(0, j, k)] = A[IDX(NX - 6, j, k)];               \
        A[IDX(1, j, k)] = A[IDX(NX - 5, j, k)];               \
        A[IDX(2, j, k)] = A[IDX(NX - 4, j, k)];               \
        A[IDX(NX - 3, j, k)] = A[IDX(3, j, k)];               \
        A[IDX(NX - 2, j, k)] = A[IDX(4, j, k)];               \
        A[IDX(NX - 1, j, k)] = A[IDX(5, j, k)];               \
    } while (0)

static void periodic_ew(orc_state *s, double *A) { /* boundaries.cpp:562-574 */
    const int NX = s->NX;
    for (int k = 0; k < s->NZ; k++)
        for (int j = 0; j < s->NY; j++) PERIODIC_EW(A);
}

/* the Fourier-space part of poisson_fft3d on s->in3d (real parts filled, levels 1..nk-2) */
static void poisson_spectral(orc_state *s) {
    int ni = s->ni, nj = s->nj, nk = s->NZ;
    cplx *tw0 = (cplx *)make_twiddles(ni, -1), *tw1 = (cplx *)make_twiddles(nj, -1);
    cplx *iw0 = (cplx *)make_twiddles(ni, +1), *iw1 = (cplx *)make_twiddles(nj, +1);
    cplx *lev1 = s->in3d + (size_t)ni * nj; /* level 1 */
    /* fftw_execute(p1), pressure.cpp:822 (levels 1..nk-2) */
    fft2d_planes(ni, nj, nk - 2, lev1, tw0, tw1);
    double size2d = (double)(ni * nj); /* pressure.cpp:824-827 */
    for (size_t n = 0; n < (size_t)ni * nj * (nk - 2); n++) { lev1[n].re /= size2d; lev1[n].im /= size2d; }
    const double *acoef = s->acoef, *ccoef = s->ccoef;
    /* forward elimination, pressure.cpp:837-856 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            double z = 1.0 / B2(i, j, 1);
            F2(i, j, 1) = ccoef[1] * z;
            IN3D(i, j, 1).re = IN3D(i, j, 1).re * z;
            IN3D(i, j, 1).im = IN3D(i, j, 1).im * z;
            for (int k = 2; k < nk - 1; k++) {
                z = 1.0 / (B2(i, j, k) - acoef[k] * F2(i, j, k - 1));
                F2(i, j, k) = ccoef[k] * z;
                IN3D(i, j, k).re = (IN3D(i, j, k).re - acoef[k] * IN3D(i, j, k - 1).re) * z;
                IN3D(i, j, k).im = (IN3D(i, j, k).im - acoef[k] * IN3D(i, j, k - 1).im) * z;
            }
        }
    /* mean-mode gauge, pressure.cpp:858-859 */
    IN3D(0, 0, nk - 2).re = 0;
    IN3D(0, 0, nk - 2).im = 0;
    /* back substitution, pressure.cpp:864-871 */
#ifdef _OPENMP
#pragma omp parallel for schedule(static)
#endif
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++)
            for (int k = nk - 3; k >= 1; k--) {
                IN3D(i, j, k).re -= F2(i, j, k) * IN3D(i, j, k + 1).re;
                IN3D(i, j, k).im -= F2(i, j, k) * IN3D(i, j, k + 1).im;
            }
    /* fftw_execute(p2), pressure.cpp:876 */
    fft2d_planes(ni, nj, nk - 2, lev1, iw0, iw1);
    free(tw0); free(tw1); free(iw0); free(iw1);
}

void orc_poisson_rhs(orc_state *s, const double *rhs, double *x) {
    int ni = s->ni, nj = s->nj, nk = s->NZ;
    for (int k = 1; k < nk - 1; k++)
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) {
                IN3D(i, j, k).re = rhs[((size_t)k * ni + i) * nj + j];
                IN3D(i, j, k).im = 0;
            }
    poisson_spectral(s);
    for (int k = 1; k < nk - 1; k++)
        for (int i = 0; i < ni; i++)
            for (int j = 0; j < nj; j++) x[((size_t)k * ni + i) * nj + j] = IN3D(i, j, k).re;
}

/* solve_pressure, serial periodic branch, pressure.cpp:202-208 */
void orc_solve_pressure(orc_state *s, double step) {
    const int NX = s->NX, NY = s->NY, NZ = s->NZ;
    int ni = s->ni, nj = s->nj;
    /* periodic_uvw_boundaries, boundaries.cpp:546-555 */
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++) {
            PERIODIC_EW(s->ups);
            PERIODIC_EW(s->vps);
            PERIODIC_EW(s->wps);
        }
    /* divergence3d_periodic(3,NX-3,1,NY-1,step), pressure.cpp:1390-1424 */
    const double *rhou = s->col[C_RHOU], *rhow = s->col[C_RHOW], *mu = s->col[C_MU];
    const double *UP = s->ups, *VP = s->vps, *WP = s->wps;
    for (int i = 3; i < NX - 3; i++)
        for (int j = 1; j < NY - 1; j++)
            for (int k = 1; k < NZ - 1; k++) {
                IN3D(i - 3, j, k).re = rhou[k] * (UP[IDX(i + 1, j, k)] - UP[IDX(i, j, k)]) * s->one_d_dx +
                                       rhou[k] * (VP[IDX(i, j + 1, k)] - VP[IDX(i, j, k)]) * s->one_d_dy +
                                       (rhow[k + 1] * WP[IDX(i, j, k + 1)] - rhow[k] * WP[IDX(i, j, k)]) * (s->one_d_dz * mu[k]);
                IN3D(i - 3, j, k).re /= (step * s->dt);
                IN3D(i - 3, j, k).im = 0; /* zeroed by the previous copy-out, pressure.cpp:882-891 */
            }
    for (int k = 1; k < NZ - 1; k++)
        for (int i = 0; i < ni; i++) {
            IN3D(i, 0, k) = IN3D(i, 1, k);
            IN3D(i, NY - 1, k) = IN3D(i, NY - 2, k);
        }
    poisson_spectral(s);
    /* copy-out, pressure.cpp:878-892 */
    for (int i = 0; i < ni; i++)
        for (int j = 0; j < nj; j++) {
            s->pis[IDX(i + 3, j, 0)] = IN3D(i, j, 1).re;
            for (int k = 1; k < NZ - 1; k++) s->pis[IDX(i + 3, j, k)] = IN3D(i, j, k).re;
            s->pis[IDX(i + 3, j, NZ - 1)] = IN3D(i, j, NZ - 2).re;
        }
    periodic_ew(s, s->pis); /* periodic_pressure_boundaries, boundaries.cpp:533-540 */
}

/* pressure-gradient force, solver.cpp:123-135 */
void orc_pgf(orc_state *s, double step, int il, int ih, int jl, int jh) {
    const double *PI = s->pis, *mw = s->col[C_MW];
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++) {
            s->ups[IDX(i, j, 1)] -= step * (s->dtx * (PI[IDX(i, j, 1)] - PI[IDX(i - 1, j, 1)]));
            s->vps[IDX(i, j, 1)] -= step * (s->dty * (PI[IDX(i, j, 1)] - PI[IDX(i, j - 1, 1)]));
            for (int k = 2; k < s->NZ - 1; k++) {
                s->ups[IDX(i, j, k)] -= step * (s->dtx * (PI[IDX(i, j, k)] - PI[IDX(i - 1, j, k)]));
                s->vps[IDX(i, j, k)] -= step * (s->dty * (PI[IDX(i, j, k)] - PI[IDX(i, j - 1, k)]));
                s->wps[IDX(i, j, k)] -= step * ((s->dtz * mw[k]) * (PI[IDX(i, j, k)] - PI[IDX(i, j, k - 1)]));
            }
        }
}

/* ------------------------------------------------------------------------------------------------
 * Boundaries (serial periodic branch), boundaries.cpp:162-191, 254-264, 304-326, 436-475
 * ----------------------------------------------------------------------------------------------*/
static void sponge_ns(orc_state *s, double *A) {
    /* sponge_boundaries_north_south: south rows [0,js), north rows [NY-jn,NY), i in [1,NX), k in [1,NZ) */
    for (int i = 1; i < s->NX; i++)
        for (int j = 0; j < s->js; j++)
            for (int k = 1; k < s->NZ; k++) A[IDX(i, j, k)] -= s->south_c[j] * A[IDX(i, j, k)];
    int nb = s->NY - s->jn;
    for (int i = 1; i < s->NX; i++)
        for (int j = nb; j < s->NY; j++)
            for (int k = 1; k < s->NZ; k++) A[IDX(i, j, k)] -= s->north_c[j - nb] * A[IDX(i, j, k)];
}
static void upper_lower(orc_state *s, double *A) { /* upper_lower_boundaries(var,1,NX,1,NY) */
    for (int i = 1; i < s->NX; i++)
        for (int j = 1; j < s->NY; j++) {
            A[IDX(i, j, 0)] = A[IDX(i, j, 1)];
            A[IDX(i, j, s->NZ - 1)] = A[IDX(i, j, s->NZ - 2)];
        }
}
static void upper_lower_zero(orc_state *s, double *A) { /* upper_lower_boundaries_zero(var,1,NX,1,NY) */
    for (int i = 1; i < s->NX; i++)
        for (int j = 1; j < s->NY; j++) {
            A[IDX(i, j, 0)] = 0;
            A[IDX(i, j, 1)] = 0;
            A[IDX(i, j, s->NZ - 1)] = 0;
        }
}

void orc_apply_bc(orc_state *s) {
    sponge_ns(s, s->ups);
    sponge_ns(s, s->vps);
    sponge_ns(s, s->thps);
    periodic_ew(s, s->ups);
    periodic_ew(s, s->vps);
    periodic_ew(s, s->pis);
    periodic_ew(s, s->thps);
    upper_lower(s, s->ups);
    upper_lower(s, s->vps);
    upper_lower(s, s->thps);
    sponge_ns(s, s->wps);
    periodic_ew(s, s->wps);
    upper_lower_zero(s, s->wps);
}

void orc_apply_bc_micro(orc_state *s) {
    sponge_ns(s, s->qvps);
    sponge_ns(s, s->qcps);
    sponge_ns(s, s->qrps);
    periodic_ew(s, s->qvps);
    periodic_ew(s, s->qcps);
    periodic_ew(s, s->qrps);
    upper_lower(s, s->qvps);
    upper_lower(s, s->qcps);
    upper_lower(s, s->qrps);
}

/* ------------------------------------------------------------------------------------------------
 * Kessler warm rain, kessler.cpp:25-213, and precip_rate, microphysics.cpp:143-160
 * ----------------------------------------------------------------------------------------------*/
static int kessler_point(double dt, double rhou, double tbk, double qbk, double tbvk, double thbar, double qbar, double pbar,
                         double pi, double *TH, double *QV, double *QC, double *QR, unsigned char *sat);
int orc_kessler_point(double dt, double rhou, double tbk, double qbk, double tbvk, double thbar, double qbar, double pbar,
                      double pi, double *TH, double *QV, double *QC, double *QR) {
    return kessler_point(dt, rhou, tbk, qbk, tbvk, thbar, qbar, pbar, pi, TH, QV, QC, QR, NULL);
}
/* sat (when not NULL) receives isSaturated of kessler.cpp:129-150, set where USE_TURBULENT_STRESS */
static int kessler_point(double dt, double rhou, double tbk, double qbk, double tbvk, double thbar, double qbar, double pbar,
                         double pi, double *TH, double *QV, double *QC, double *QR, unsigned char *sat) {
    const double cpRd = cp / Rd;
    double A, B, C, E, cvent, qvsat, pd, theta, pressure, phi, vapor, total_convert, temperature, dc, de, esl, f;
    int err = 0;
    theta = *TH + thbar + tbk;
    pressure = pi / (cp * tbvk) + pbar;
    vapor = *QV + qbar + qbk;
    temperature = theta * pressure;
    pd = p0 * pow(pressure, cpRd);
    if (*QC > 1.0e-3) A = 0.001 * (*QC - 1.0e-3) * dt;
    else A = 0;
    if (*QR > minvar) B = rhou * 2.2 * *QC * pow(*QR, 0.875) * dt;
    else B = 0;
    if (*QR > minvar) {
        esl = 611.2 * exp(17.67 * (temperature - 273.15) / (temperature - 29.65));
        qvsat = 0.62197 * esl / (pd - esl);
        cvent = 1.6 + 30.3922 * pow(rhou * *QR, 0.2046);
        E = (1.0 / rhou) * (((1.0 - vapor / qvsat) * cvent * pow(rhou * *QR, 0.525)) / (2.03e4 + 9.58e6 / (pd * qvsat))) * dt;
        E = fmin(*QR, E);
        E = fmax(E, 0.0);
        if (E > 100) err = 1;
    } else E = 0;
    *QV += E;
    total_convert = fmin(A + B, *QC);
    *QC -= total_convert;
    *QR += (total_convert - E);
    de = (Lv / (cp * pressure)) * E;
    *TH -= de;
    theta = *TH + thbar + tbk;
    vapor = *QV + qbar + qbk;
    temperature = theta * pressure;
    esl = 611.2 * exp(17.67 * (temperature - 273.15) / (temperature - 29.65));
    qvsat = 0.62197 * esl / (pd - esl);
    f = 17.67 * (273.15 - 29.65) * Lv / cp;
    phi = pd / (pd - esl) * qvsat * f / pow((temperature - 29.65), 2);
    C = 0.0;
    if (vapor > qvsat) {
        C = (vapor - qvsat) / (1.0 + phi);
        if (sat) *sat = 1;
    } else if (vapor < qvsat && *QC > 0) {
        C = -fmin(*QC, (qvsat - vapor) / (1.0 + phi));
        if (sat) *sat = (-C < *QC) ? 1 : 0;
    } else if (sat) {
        *sat = (vapor == qvsat) ? 1 : 0;
    }
    *QC += C;
    *QV -= C;
    dc = (Lv / (cp * pressure)) * C;
    *TH += dc;
    return err;
}

void orc_kessler(orc_state *s, int il, int ih, int jl, int jh) {
    const double *rhou = s->col[C_RHOU], *tb = s->col[C_TB], *qb = s->col[C_QB], *tbv = s->col[C_TBV];
    int err = 0;
#ifdef _OPENMP
#pragma omp parallel for schedule(static) reduction(| : err)
#endif
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int k = 1; k < s->NZ - 1; k++) {
                size_t n = IDX(i, j, k);
                err |= kessler_point(s->dt, rhou[k], tb[k], qb[k], tbv[k], s->thbar[n], s->qbar[n], s->pbar[n], s->pis[n], &s->thps[n],
                                     &s->qvps[n], &s->qcps[n], &s->qrps[n], s->turb ? &s->saturated[n] : NULL);
            }
    if (err) s->error = 1;
    /* precip_rate(il,ih,jl,jh,vts,qrps,accRain), kessler.cpp:210 -> microphysics.cpp:143-160 */
    const double *rhow = s->col[C_RHOW];
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            s->accrain[(size_t)i * s->NY + j] += rhow[1] * 0.5 * (s->vts[IDX(i, j, 0)] + s->vts[IDX(i, j, 1)]) * 0.5 *
                                                 (s->qrps[IDX(i, j, 0)] + s->qrps[IDX(i, j, 1)]) * s->dt;
}

/* rayleigh_damping, damping.cpp:1148-1177 (unstretched zu, zw) */
void orc_rayleigh(orc_state *s, int il, int ih, int jl, int jh, int level) {
    if (level >= s->NZ) return; /* the reference reads zu[NZ] here and damps nothing */
    const double *zu = s->col[C_ZU], *zw = s->col[C_ZW];
    double raydmpcoef = 0.05, raydmpz = zu[level];
    for (int k = level; k < s->NZ; k++) {
        double coefu = raydmpcoef * 0.5 * (1.0 - cos(trigpi * (zu[k] - raydmpz) / (zu[s->NZ - 1] - raydmpz)));
        double coefw = raydmpcoef * 0.5 * (1.0 - cos(trigpi * (zw[k] - raydmpz) / (zw[s->NZ - 1] - raydmpz)));
        for (int i = il; i < ih; i++)
            for (int j = jl; j < jh; j++) {
                size_t n = IDX(i, j, k);
                s->ups[n] -= coefu * s->ups[n];
                s->vps[n] -= coefu * s->vps[n];
                s->wps[n] -= coefw * s->wps[n];
                s->thps[n] -= coefu * s->thps[n];
            }
    }
}

/* ------------------------------------------------------------------------------------------------
 * Turbulence parameterization, turbulence.cpp (USE_TURBULENT_STRESS, no terrain: HTOPO = 0, no surface
 * heat flux: SURFACE_HEAT_FLUX 0).  Mixing coefficients at the start of each step from the M buffers,
 * tendencies stored and added during every stage.
 * ----------------------------------------------------------------------------------------------*/
static const double cmixv_max = 0.15, cmixh_max = 0.05; /* turbulence.cpp:48-49 */
static const double one_d_Pr = 3.0, lv_max = 200.0, cs_turb = 0.25; /* :56-58 */

#define KM(i, j, k) s->kmix[IDX(i, j, k)]
#define KHM(i, j, k) s->khmix[IDX(i, j, k)]
#define ONE_D_DZ(k) (s->one_d_dz * s->col[C_MU][k])
#define ONE_D_DZW(k) (s->one_d_dz * s->col[C_MW][k])

/* init_kmix, turbulence.cpp:85-147, with zlevs = ZU = zsu (initialize_serial, driver_serial.cpp:52) */
static void turb_init(orc_state *s) {
    s->kmixh_max = cmixh_max * s->dx * s->dx / s->dt;
    s->kmixv_max = cmixv_max * s->dz * s->dz / s->dt;
    s->lH = s->dx / 4.0;
    const double *z = s->col[C_ZSU];
    for (int k = 1; k < s->NZ; k++) {
        double deltaZ = z[k] - z[k - 1];
        s->vml2[k] = deltaZ > lv_max ? lv_max * lv_max : deltaZ * deltaZ;
        s->vkmax[k] = cmixv_max * s->vml2[k] / s->dt;
    }
}

/* the six rate-of-deformation products, get_Kmix(_dry), turbulence.cpp:596-624 */
static double deformation2(const orc_state *s, int i, int j, int k) {
    const double *UM = s->ums, *VM = s->vms, *WM = s->wms, *UB = s->ubar, *VB = s->vbar, *WB = s->wbar;
    const double odx = s->one_d_dx, ody = s->one_d_dy;
#define A3(F, a, b, c) F[IDX(a, b, c)]
    double D11 = 0.50 * (A3(UM, i + 1, j, k) + A3(UM, i + 1, j, k - 1) - A3(UM, i - 1, j, k) - A3(UM, i - 1, j, k - 1) +
                         A3(UB, i + 1, j, k) + A3(UB, i + 1, j, k - 1) - A3(UB, i - 1, j, k) - A3(UB, i - 1, j, k - 1)) * odx;
    double D22 = 0.50 * (A3(VM, i, j + 1, k) + A3(VM, i, j + 1, k - 1) - A3(VM, i, j - 1, k) - A3(VM, i, j - 1, k - 1) +
                         A3(VB, i, j + 1, k) + A3(VB, i, j + 1, k - 1) - A3(VB, i, j - 1, k) - A3(VB, i, j - 1, k - 1)) * ody;
    double D33 = 1.0 * (A3(WM, i, j, k + 1) - A3(WM, i, j, k - 1) + A3(WB, i, j, k + 1) - A3(WB, i, j, k - 1)) * ONE_D_DZW(k);
    double D12 = 0.5 * ((A3(UM, i, j + 1, k) + A3(UM, i, j + 1, k - 1) - A3(UM, i, j - 1, k) - A3(UM, i, j - 1, k - 1) +
                         A3(UB, i, j + 1, k) + A3(UB, i, j + 1, k - 1) - A3(UB, i, j - 1, k) - A3(UB, i, j - 1, k - 1)) * ody +
                        (A3(VM, i + 1, j, k) + A3(VM, i + 1, j, k - 1) - A3(VM, i - 1, j, k) - A3(VM, i - 1, j, k - 1) +
                         A3(VB, i + 1, j, k) + A3(VB, i + 1, j, k - 1) - A3(VB, i - 1, j, k) - A3(VB, i - 1, j, k - 1)) * odx);
    double D13 = 0.5 * ((A3(UM, i, j, k) + A3(UM, i + 1, j, k) - A3(UM, i, j, k - 1) - A3(UM, i + 1, j, k - 1) +
                         A3(UB, i, j, k) + A3(UB, i + 1, j, k) - A3(UB, i, j, k - 1) - A3(UB, i + 1, j, k - 1)) * ONE_D_DZW(k) +
                        (A3(WM, i + 1, j, k) - A3(WM, i - 1, j, k) + A3(WB, i + 1, j, k) - A3(WB, i - 1, j, k)) * odx);
    double D23 = 0.5 * ((A3(VM, i, j, k) + A3(VM, i, j + 1, k) - A3(VM, i, j, k - 1) - A3(VM, i, j + 1, k - 1) +
                         A3(VB, i, j, k) + A3(VB, i, j + 1, k) - A3(VB, i, j, k - 1) - A3(VB, i, j + 1, k - 1)) * ONE_D_DZW(k) +
                        (A3(WM, i, j + 1, k) - A3(WM, i, j - 1, k) + A3(WB, i, j + 1, k) - A3(WB, i, j - 1, k)) * ody);
#undef A3
    return D11 * D11 + D22 * D22 + D33 * D33 + D12 * D12 + D13 * D13 + D23 * D23;
}

/* the Brunt-Vaisala frequency of get_Kmix_dry (turbulence.cpp:781-791) and get_Kmix (:627-690) */
static double brunt_vaisala(const orc_state *s, int i, int j, int k) {
    const double *tb = s->col[C_TB], *qb = s->col[C_QB], *tbv = s->col[C_TBV];
    const size_t h = IDX(i, j, k), l = IDX(i, j, k - 1);
    if (!s->moist || !s->saturated[h] || !s->saturated[l]) {
        double qh = s->moist ? s->qvms[h] + s->qbar[h] + qb[k] : qb[k];
        double ql = s->moist ? s->qvms[l] + s->qbar[l] + qb[k - 1] : qb[k - 1];
        double thetaVH = (s->thms[h] + s->thbar[h] + tb[k]) * (1 + 0.61 * qh);
        double thetaVL = (s->thms[l] + s->thbar[l] + tb[k - 1]) * (1 + 0.61 * ql);
        double dTheta = thetaVH - thetaVL;
        return 2.0 * grav / (tbv[k] + tbv[k - 1]) * dTheta * ONE_D_DZW(k);
    }
    /* saturated: equivalent potential temperature and the moist-adiabatic lapse (turbulence.cpp:652-690) */
    const double *rhou = s->col[C_RHOU], *tbw = s->col[C_TBW];
    const double cpRd = cp / Rd;
    double thetaVH = s->thms[h] + s->thbar[h] + tb[k];
    double thetaVL = s->thms[l] + s->thbar[l] + tb[k - 1];
    double tempH = thetaVH * (s->pbar[h] + s->pis[h] / (cp * tbv[k]));
    double tempL = thetaVL * (s->pbar[l] + s->pis[l] / (cp * tbv[k - 1]));
    double temp = 0.5 * (tempH + tempL);
    double dQL = s->qcms[h] + s->qrms[h] - s->qcms[l] - s->qrms[l] + s->qvms[h] - s->qvms[l] + s->qbar[h] - s->qbar[l] + qb[k] - qb[k - 1];
    double qv = 0.5 * (s->qvms[h] + s->qvms[l] + s->qbar[h] + s->qbar[l] + qb[k] + qb[k - 1]);
    double pdH = p0 * pow(s->pbar[h], cpRd) + s->pis[h] * rhou[k];
    double pdL = p0 * pow(s->pbar[l], cpRd) + s->pis[l] * rhou[k - 1];
    double eslH = 611.2 * exp(17.67 * (tempH - 273.15) / (tempH - 29.65));
    double eslL = 611.2 * exp(17.67 * (tempL - 273.15) / (tempL - 29.65));
    double qvsatH = 0.62197 * eslH / (pdH - eslH);
    double qvsatL = 0.62197 * eslL / (pdL - eslL);
    thetaVH = thetaVH * (1 + Lv * qvsatH / (cp * tempH));
    thetaVL = thetaVL * (1 + Lv * qvsatL / (cp * tempL));
    double dTheta = thetaVH - thetaVL;
    double A = (grav / tbw[k]) * (1.0 + Lv * qv / (Rd * temp)) / (1.0 + 0.622 * Lv * Lv * qv / (cp * Rd * temp * temp));
    return (A * dTheta - grav * dQL) * ONE_D_DZW(k);
}

/* get_Kmix / get_Kmix_dry, turbulence.cpp:563-735, 737-831: vertical mixing coefficient on w-levels,
 * horizontal on u-levels (level k-1 set with level k), both bounded above */
static void turb_kmix(orc_state *s, int il, int ih, int jl, int jh) {
    const int NZ = s->NZ;
    const double lH2 = s->lH * s->lH;
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++) {
            KM(i, j, 1) = 0; /* zero at ground */
            KM(i, j, NZ - 1) = 0;
            for (int k = 2; k < NZ - 1; k++) {
                double S2 = deformation2(s, i, j, k);
                double bruntv = brunt_vaisala(s, i, j, k);
                if (S2 - one_d_Pr * bruntv > 0) {
                    double c = s->vml2[k] * cs_turb * cs_turb;
                    KM(i, j, k) = c * sqrt(S2 - one_d_Pr * bruntv);
                    KHM(i, j, k - 1) = 0.5 * (KM(i, j, k - 1) / s->vml2[k - 1] + KM(i, j, k) / s->vml2[k]) * s->lH * s->lH;
                    if (KM(i, j, k) > s->vkmax[k]) KM(i, j, k) = s->vkmax[k];
                    if (KHM(i, j, k - 1) > s->kmixh_max) KHM(i, j, k - 1) = s->kmixh_max;
                } else {
                    KM(i, j, k) = 0;
                    KHM(i, j, k - 1) = 0 + 0.5 * KM(i, j, k - 1) / s->vml2[k - 1] * s->lH * s->lH;
                }
            }
        }
    (void)lH2;
}

/* components of the rate-of-deformation tensor at slab i (turbulence.cpp:913-933); level 0 of tau_33,
 * tau_13 and tau_23 is never computed and stays zero (the stress cells are zero-filled) */
static double tau11(const orc_state *s, int i, int j, int k) { return 2.0 * (s->ums[IDX(i + 1, j, k)] - s->ums[IDX(i, j, k)]) * s->one_d_dx; }
static double tau22(const orc_state *s, int i, int j, int k) { return 2.0 * (s->vms[IDX(i, j + 1, k)] - s->vms[IDX(i, j, k)]) * s->one_d_dy; }
static double tau33(const orc_state *s, int i, int j, int k) {
    return k < 1 ? 0.0 : 2.0 * (s->wms[IDX(i, j, k + 1)] - s->wms[IDX(i, j, k)]) * ONE_D_DZ(k);
}
static double tau13(const orc_state *s, int i, int j, int k) {
    if (k < 1) return 0.0;
    return (s->wms[IDX(i + 1, j, k + 1)] - s->wms[IDX(i, j, k + 1)]) * s->one_d_dx +
           (s->ums[IDX(i + 1, j, k + 1)] - s->ums[IDX(i + 1, j, k)]) * ONE_D_DZW(k + 1);
}
static double tau23(const orc_state *s, int i, int j, int k) {
    if (k < 1) return 0.0;
    return (s->wms[IDX(i, j + 1, k + 1)] - s->wms[IDX(i, j, k + 1)]) * s->one_d_dy +
           (s->vms[IDX(i, j + 1, k + 1)] - s->vms[IDX(i, j + 1, k)]) * ONE_D_DZW(k + 1);
}
static double tau12(const orc_state *s, int i, int j, int k) {
    return (s->ums[IDX(i + 1, j + 1, k)] - s->ums[IDX(i + 1, j, k)]) * s->one_d_dy +
           (s->vms[IDX(i + 1, j + 1, k)] - s->vms[IDX(i, j + 1, k)]) * s->one_d_dx;
}

/* surface stress of the lowest level (turbulence.cpp:885-907): land (landsea < 0.5) a fixed drag
 * coefficient, ocean one that grows with the wind speed */
static void surface_stress(const orc_state *s, int i, int j, double *tu, double *tv) {
    const size_t n = IDX(i, j, 1);
    double wind_speed = sqrt((s->ums[n] + s->ubar[n]) * (s->ums[n] + s->ubar[n]) + (s->vms[n] + s->vbar[n]) * (s->vms[n] + s->vbar[n]));
    double wind_speed_base = sqrt(s->ubar[n] * s->ubar[n] + s->vbar[n] * s->vbar[n]);
    if (s->landsea[(size_t)i * s->NY + j] < 0.5) {
        *tu = 0.003 * (wind_speed * s->ums[n] + (wind_speed - wind_speed_base) * s->ubar[n]);
        *tv = 0.003 * (wind_speed * s->vms[n] + (wind_speed - wind_speed_base) * s->vbar[n]);
    } else {
        double drag_coef = (1.1e-3 + 4.0e-5 * wind_speed);
        *tu = drag_coef * (wind_speed * s->ums[n] + (wind_speed - wind_speed_base) * s->ubar[n]);
        *tv = drag_coef * (wind_speed * s->vms[n] + (wind_speed - wind_speed_base) * s->vbar[n]);
    }
}

/* turbulent_diffusion_velocity, turbulence.cpp:839-1020: friction tendencies of u, v, w over
 * [il,ih) x [jl,jh) x [1,NZ-1); the stress of slab i-1 is the west stress of slab i */
static void turb_velocity(orc_state *s, int il, int ih, int jl, int jh) {
    const int NZ = s->NZ;
    const double odx = s->one_d_dx, ody = s->one_d_dy;
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++) {
            double tau_u_zl, tau_v_zl;
            surface_stress(s, i, j, &tau_u_zl, &tau_v_zl);
            for (int k = 1; k < NZ - 1; k++) {
                double uK11H_h = KHM(i, j, k);
                double uK11H_l = KHM(i - 1, j, k);
                double uK12H_h = 0.25 * (KHM(i, j, k) + KHM(i - 1, j, k) + KHM(i - 1, j + 1, k) + KHM(i, j + 1, k));
                double uK12H_l = 0.25 * (KHM(i, j - 1, k) + KHM(i - 1, j - 1, k) + KHM(i - 1, j, k) + KHM(i, j, k));
                double uK13V_h = 0.50 * (KM(i, j, k + 1) + KM(i - 1, j, k + 1));
                double vK12H_h = 0.25 * (KHM(i, j - 1, k) + KHM(i, j, k) + KHM(i + 1, j - 1, k) + KHM(i + 1, j, k));
                double vK12H_l = uK12H_l;
                double vK22H_h = KHM(i, j, k);
                double vK22H_l = KHM(i, j - 1, k);
                double vK23V_h = 0.5 * (KM(i, j, k + 1) + KM(i, j - 1, k + 1));
                double wK13H_h = 0.25 * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i + 1, j, k) + KHM(i + 1, j, k - 1));
                double wK13H_l = 0.25 * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i - 1, j, k) + KHM(i - 1, j, k - 1));
                double wK23H_h = 0.25 * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i, j + 1, k) + KHM(i, j + 1, k - 1));
                double wK23H_l = 0.25 * (KHM(i, j, k) + KHM(i, j, k - 1) + KHM(i, j - 1, k) + KHM(i, j - 1, k - 1));
                double wK33V_h = 0.50 * (KM(i, j, k) + KM(i, j, k + 1));
                double wK33V_l = 0.50 * (KM(i, j, k) + KM(i, j, k - 1));
                double ufric = (uK11H_h * tau11(s, i, j, k) - uK11H_l * tau11(s, i - 1, j, k)) * odx +
                               (uK12H_h * tau12(s, i - 1, j, k) - uK12H_l * tau12(s, i - 1, j - 1, k)) * ody +
                               (uK13V_h * tau13(s, i - 1, j, k) - tau_u_zl) * ONE_D_DZ(k);
                double vfric = (vK12H_h * tau12(s, i, j - 1, k) - vK12H_l * tau12(s, i - 1, j - 1, k)) * odx +
                               (vK22H_h * tau22(s, i, j, k) - vK22H_l * tau22(s, i, j - 1, k)) * ody +
                               (vK23V_h * tau23(s, i, j - 1, k) - tau_v_zl) * ONE_D_DZ(k);
                double wfric = (wK13H_h * tau13(s, i, j, k - 1) - wK13H_l * tau13(s, i - 1, j, k - 1)) * odx +
                               (wK23H_h * tau23(s, i, j, k - 1) - wK23H_l * tau23(s, i, j - 1, k - 1)) * ody +
                               (wK33V_h * tau33(s, i, j, k) - wK33V_l * tau33(s, i, j, k - 1)) * ONE_D_DZW(k);
                s->ufric[IDX(i, j, k)] = ufric;
                s->vfric[IDX(i, j, k)] = vfric;
                s->wfric[IDX(i, j, k)] = wfric;
                tau_u_zl = uK13V_h * tau13(s, i - 1, j, k);
                tau_v_zl = vK23V_h * tau23(s, i, j - 1, k);
            }
        }
}

/* the diffusion tendency of one scalar at (i,j,k) (turbulence.cpp:1128-1174, 1226-1266); the bottom
 * flux is the top flux of the level below, zero at the ground (no surface heat flux) */
static double scalar_diffusion(const orc_state *s, const double *S, int i, int j, int k, double tau_l) {
    double tau_h = -one_d_Pr * KM(i, j, k + 1) * (S[IDX(i, j, k + 1)] - S[IDX(i, j, k)]) * ONE_D_DZW(k + 1);
    double KH = 0.5 * (one_d_Pr * KHM(i, j + 1, k) + one_d_Pr * KHM(i, j, k));
    double KL = 0.5 * (one_d_Pr * KHM(i, j, k) + one_d_Pr * KHM(i, j - 1, k));
    double t22h = -KH * (S[IDX(i, j + 1, k)] - S[IDX(i, j, k)]) * s->one_d_dy;
    double t22l = -KL * (S[IDX(i, j, k)] - S[IDX(i, j - 1, k)]) * s->one_d_dy;
    KH = 0.5 * (one_d_Pr * KHM(i + 1, j, k) + one_d_Pr * KHM(i, j, k));
    KL = 0.5 * (one_d_Pr * KHM(i, j, k) + one_d_Pr * KHM(i - 1, j, k));
    double t11h = -KH * (S[IDX(i + 1, j, k)] - S[IDX(i, j, k)]) * s->one_d_dx;
    double t11l = -KL * (S[IDX(i, j, k)] - S[IDX(i - 1, j, k)]) * s->one_d_dx;
    return -((t11h - t11l) * s->one_d_dx + (t22h - t22l) * s->one_d_dy + (tau_h - tau_l) * ONE_D_DZ(k));
}

/* turbulent_diffusion_theta / turbulent_diffusion_scalars, turbulence.cpp:1027-1188, 1192-1276 */
static void turb_scalars(orc_state *s, int il, int ih, int jl, int jh) {
    const int NZ = s->NZ;
    const double *F[4] = {s->thms, s->qvms, s->qcms, s->qrms};
    double *T[4] = {s->tdiff, s->qvdiff, s->qcdiff, s->qrdiff};
    const int nf = s->moist ? 4 : 1;
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int f = 0; f < nf; f++) {
                double tau_l = 0.0;
                for (int k = 1; k < NZ - 1; k++) {
                    T[f][IDX(i, j, k)] = scalar_diffusion(s, F[f], i, j, k, tau_l);
                    tau_l = -one_d_Pr * KM(i, j, k + 1) * (F[f][IDX(i, j, k + 1)] - F[f][IDX(i, j, k)]) * ONE_D_DZW(k + 1);
                }
            }
}

void orc_calculate_diff_tend(orc_state *s, int il, int ih, int jl, int jh) {
    if (!s->turb) return;
    turb_kmix(s, il, ih, jl, jh);
    turb_velocity(s, il, ih, jl, jh);
    turb_scalars(s, il, ih, jl, jh);
}

/* apply_velocity_diffusion, turbulence.cpp:233-258: w from level 2 */
void orc_apply_velocity_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh) {
    double deltaT = step * s->dt;
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++) {
            size_t ind = IDX(i, j, 1);
            s->ups[ind] += s->ufric[ind] * deltaT;
            s->vps[ind] += s->vfric[ind] * deltaT;
            for (int k = 2; k < s->NZ - 1; k++) {
                ind = IDX(i, j, k);
                s->ups[ind] += s->ufric[ind] * deltaT;
                s->vps[ind] += s->vfric[ind] * deltaT;
                s->wps[ind] += s->wfric[ind] * deltaT;
            }
        }
}

/* apply_scalar_diffusion, turbulence.cpp:263-273 */
void orc_apply_scalar_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh) {
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int k = 1; k < s->NZ - 1; k++) s->thps[IDX(i, j, k)] += s->tdiff[IDX(i, j, k)] * step * s->dt;
}

/* apply_moisture_diffusion, turbulence.cpp:278-292 */
void orc_apply_moisture_diffusion(orc_state *s, double step, int il, int ih, int jl, int jh) {
    double deltaT = step * s->dt;
    for (int i = il; i < ih; i++)
        for (int j = jl; j < jh; j++)
            for (int k = 1; k < s->NZ - 1; k++) {
                size_t n = IDX(i, j, k);
                s->qvps[n] += s->qvdiff[n] * deltaT;
                s->qcps[n] += s->qcdiff[n] * deltaT;
                s->qrps[n] += s->qrdiff[n] * deltaT;
            }
}

/* USE_TURBULENT_STRESS with the land-sea mask (NULL: all ocean, initialize_landsea without a file,
 * turbulence.cpp:370-375) */
void orc_set_turbulence(orc_state *s, int on, const double *landsea) {
    s->turb = on;
    if (!on) return;
    size_t size = (size_t)s->NX * s->NYNZ;
    double **bufs[] = {&s->kmix, &s->khmix, &s->ufric, &s->vfric, &s->wfric, &s->tdiff, &s->qvdiff, &s->qcdiff, &s->qrdiff};
    for (size_t b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++)
        if (!*bufs[b]) *bufs[b] = (double *)calloc(size, sizeof(double));
    if (!s->saturated) s->saturated = (unsigned char *)calloc(size, 1);
    if (!s->landsea) s->landsea = (double *)calloc((size_t)s->NX * s->NY, sizeof(double));
    for (size_t n = 0; n < (size_t)s->NX * s->NY; n++) s->landsea[n] = landsea ? landsea[n] : 1.0;
    if (!s->vml2) s->vml2 = (double *)calloc(s->NZ, sizeof(double));
    if (!s->vkmax) s->vkmax = (double *)calloc(s->NZ, sizeof(double));
    turb_init(s);
}

/* the turbulence arrays, for tests: 0 Kmix, 1 KHmix, 2-4 u/v/w friction, 5-8 theta/qv/qc/qr diffusion,
 * 9 Kessler's saturation flag (0 or 1) */
void orc_get_turbulence(const orc_state *s, int which, double *out) {
    const double *src[] = {s->kmix, s->khmix, s->ufric, s->vfric, s->wfric, s->tdiff, s->qvdiff, s->qcdiff, s->qrdiff};
    if (which == 9 && s->saturated) {
        for (size_t n = 0; n < (size_t)s->NX * s->NYNZ; n++) out[n] = s->saturated[n];
        return;
    }
    if (which < 0 || which > 8 || !src[which]) return;
    memcpy(out, src[which], sizeof(double) * (size_t)s->NX * s->NYNZ);
}

/* Kessler's saturation flags set directly, one byte per point [NX][NY][NZ] (turbulence on), so tests can
 * reach the saturated branch of the Brunt-Vaisala frequency with states of their choosing */
void orc_set_saturated(orc_state *s, const unsigned char *flags) {
    if (!s->saturated) return;
    memcpy(s->saturated, flags, (size_t)s->NX * s->NYNZ);
}

/* ------------------------------------------------------------------------------------------------
 * Buffer rotation, solver.cpp:229-263 and microphysics.cpp:642-669
 * ----------------------------------------------------------------------------------------------*/
static void swap(double **a, double **b) { double *t = *a; *a = *b; *b = t; }

void orc_advance_inner(orc_state *s) {
    swap(&s->us, &s->ups);
    swap(&s->vs, &s->vps);
    swap(&s->ths, &s->thps);
    swap(&s->ws, &s->wps);
}
void orc_advance_outer(orc_state *s) {
    size_t bytes = sizeof(double) * (size_t)s->NX * s->NYNZ;
    memcpy(s->ums, s->ups, bytes);
    memcpy(s->vms, s->vps, bytes);
    memcpy(s->thms, s->thps, bytes);
    memcpy(s->wms, s->wps, bytes);
    orc_advance_inner(s);
}
void orc_micro_advance_inner(orc_state *s) {
    swap(&s->qvs, &s->qvps);
    swap(&s->qcs, &s->qcps);
    swap(&s->qrs, &s->qrps);
}
void orc_micro_advance_outer(orc_state *s) {
    size_t bytes = sizeof(double) * (size_t)s->NX * s->NYNZ;
    memcpy(s->qvms, s->qvps, bytes);
    memcpy(s->qcms, s->qcps, bytes);
    memcpy(s->qrms, s->qrps, bytes);
    orc_micro_advance_inner(s);
}

/* ------------------------------------------------------------------------------------------------
 * s_integrate_rk3, driver_serial.cpp:178-286 (no explicit diffusion, no terrain; the prescribed heating
 * never fires, Heating.h:3; turbulence when orc_set_turbulence turned it on)
 * ----------------------------------------------------------------------------------------------*/
static void integrate_rk3(orc_state *s) {
    const double steps[] = {1. / 3., 0.5, 1.0};
    const int NX = s->NX, NY = s->NY;
    if (s->turb) orc_calculate_diff_tend(s, 1, NX - 1, 1, NY - 1); /* driver_serial.cpp:194 */
    for (int st = 0; st < 3; st++) {
        double step = steps[st];
        /* integrate_non_hydro(step,0,NX,0,NY), solver.cpp:87-138 */
        orc_advect_uvw(s, step, 3, NX - 3, 3, NY - 3);
        if (s->turb) orc_apply_velocity_diffusion(s, step, 3, NX - 3, 3, NY - 3); /* solver.cpp:104 */
        orc_solve_pressure(s, step);
        orc_pgf(s, step, 1, NX - 1, 1, NY - 1);
        orc_advect_theta(s, step, 3, NX - 3, 3, NY - 3);
        if (s->turb) orc_apply_scalar_diffusion(s, step, 1, NX - 1, 1, NY - 1); /* driver_serial.cpp:218 */
        if (s->moist) {
            orc_advect_moisture(s, step, 3, NX - 3, 3, NY - 3);
            if (s->turb) orc_apply_moisture_diffusion(s, step, 1, NX - 1, 1, NY - 1); /* driver_serial.cpp:231 */
            orc_zero_moisture(s, 1, NX - 1, 1, NY - 1);
        }
        orc_apply_bc(s);
        if (s->moist) {
            orc_apply_bc_micro(s);
            if (st < 2) orc_micro_advance_inner(s);
        }
        if (st < 2) orc_advance_inner(s);
    }
    if (s->moist) {
        orc_kessler(s, 1, NX - 1, 1, NY - 1);
        orc_apply_bc_micro(s);
        orc_micro_advance_outer(s);
    }
    orc_apply_bc(s);
    orc_rayleigh(s, 1, NX - 1, 1, NY - 1, s->raylevel);
    orc_advance_outer(s);
}

void orc_step(orc_state *s, int nsteps) {
    for (int n = 0; n < nsteps; n++) integrate_rk3(s);
}

/* ------------------------------------------------------------------------------------------------
 * state management
 * ----------------------------------------------------------------------------------------------*/
/* init_boundaries sponge coefficients via cos_profile, boundaries.cpp:84-86, util.cpp:740-749 */
static double cos_profile(double pl, double ph, double frac, double shift) {
    if (frac >= 0 && frac <= 1.0) return cos(pl * 2 * trigpi + (ph - pl) * 2 * trigpi * frac) + shift;
    return 0;
}

orc_state *orc_create(const orc_grid *g) {
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->NX = g->nx; s->NY = g->ny; s->NZ = g->nz; s->NYNZ = g->ny * g->nz;
    s->dx = g->dx; s->dy = g->dx; s->dz = g->dz; s->dt = g->dt;
    s->dtx = s->dt / s->dx; s->dty = s->dt / s->dy; s->dtz = s->dt / s->dz; /* initializer.cpp:173-179 */
    s->one_d_dx = 1.0 / s->dx; s->one_d_dy = 1.0 / s->dy; s->one_d_dz = 1.0 / s->dz;
    s->moist = g->microphysics; s->raylevel = g->rayleigh_level; s->jn = g->sponge_north; s->js = g->sponge_south;
    for (int c = 0; c < C_COUNT; c++) s->col[c] = (double *)calloc(c == C_F ? s->NY : s->NZ, sizeof(double));
    size_t size = (size_t)s->NX * s->NYNZ;
    double **bufs[] = {&s->us, &s->ups, &s->ums, &s->vs, &s->vps, &s->vms, &s->ws, &s->wps, &s->wms, &s->ths, &s->thps, &s->thms,
                       &s->qvs, &s->qvps, &s->qvms, &s->qcs, &s->qcps, &s->qcms, &s->qrs, &s->qrps, &s->qrms, &s->pis, &s->vts,
                       &s->ubar, &s->vbar, &s->wbar, &s->thbar, &s->qbar, &s->pbar};
    for (size_t b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) *bufs[b] = (double *)calloc(size, sizeof(double));
    s->accrain = (double *)calloc((size_t)s->NX * s->NY, sizeof(double));
    s->north_c = (double *)calloc(s->jn > 0 ? s->jn : 1, sizeof(double));
    s->south_c = (double *)calloc(s->js > 0 ? s->js : 1, sizeof(double));
    for (int i = 0; i < s->jn; i++) s->north_c[i] = 0.5 * cos_profile(-0.5, 0, (double)(i + 1) / (double)s->jn, 1);
    for (int i = 0; i < s->js; i++) s->south_c[i] = 0.5 * cos_profile(0, 0.5, (double)i / (double)s->js, 1);
    init_pressure_solver(s);
    return s;
}

void orc_destroy(orc_state *s) {
    if (!s) return;
    for (int c = 0; c < C_COUNT; c++) free(s->col[c]);
    double *bufs[] = {s->us, s->ups, s->ums, s->vs, s->vps, s->vms, s->ws, s->wps, s->wms, s->ths, s->thps, s->thms,
                      s->qvs, s->qvps, s->qvms, s->qcs, s->qcps, s->qcms, s->qrs, s->qrps, s->qrms, s->pis, s->vts,
                      s->ubar, s->vbar, s->wbar, s->thbar, s->qbar, s->pbar, s->accrain, s->north_c, s->south_c,
                      s->kwave, s->lwave, s->acoef, s->ccoef, s->bcoef, s->rhs,
                      s->kmix, s->khmix, s->ufric, s->vfric, s->wfric, s->tdiff, s->qvdiff, s->qcdiff, s->qrdiff,
                      s->landsea, s->vml2, s->vkmax};
    for (size_t b = 0; b < sizeof(bufs) / sizeof(bufs[0]); b++) free(bufs[b]);
    free(s->in3d);
    free(s->saturated);
    free(s);
}

int orc_error(const orc_state *s) { return s->error; }

void orc_set_column(orc_state *s, int c, const double *v) {
    memcpy(s->col[c], v, sizeof(double) * (size_t)(c == C_F ? s->NY : s->NZ));
    if (c == C_MU || c == C_MW || c == C_RHOW || c == C_RHOU) pressure_coefficients(s);
}
void orc_get_column(const orc_state *s, int c, double *v) { memcpy(v, s->col[c], sizeof(double) * (size_t)(c == C_F ? s->NY : s->NZ)); }

static double **field_ptr(orc_state *s, int f, int role) {
    switch (f) {
    case F_U: return role == R_M ? &s->ums : role == R_CUR ? &s->us : &s->ups;
    case F_V: return role == R_M ? &s->vms : role == R_CUR ? &s->vs : &s->vps;
    case F_W: return role == R_M ? &s->wms : role == R_CUR ? &s->ws : &s->wps;
    case F_TH: return role == R_M ? &s->thms : role == R_CUR ? &s->ths : &s->thps;
    case F_QV: return role == R_M ? &s->qvms : role == R_CUR ? &s->qvs : &s->qvps;
    case F_QC: return role == R_M ? &s->qcms : role == R_CUR ? &s->qcs : &s->qcps;
    case F_QR: return role == R_M ? &s->qrms : role == R_CUR ? &s->qrs : &s->qrps;
    case F_PI: return &s->pis;
    case F_UBAR: return &s->ubar;
    case F_VBAR: return &s->vbar;
    case F_WBAR: return &s->wbar;
    case F_THBAR: return &s->thbar;
    case F_QBAR: return &s->qbar;
    case F_PBAR: return &s->pbar;
    case F_VT: return &s->vts;
    case F_ACCRAIN: return &s->accrain;
    }
    return NULL;
}

void orc_set_field(orc_state *s, int f, int role, const double *v) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    memcpy(*field_ptr(s, f, role), v, sizeof(double) * n);
}
void orc_get_field(const orc_state *s, int f, int role, double *v) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    memcpy(v, *field_ptr((orc_state *)s, f, role), sizeof(double) * n);
}
