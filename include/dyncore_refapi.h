/*
 * dyncore_refapi.h -- the reference's own dynamical-core entry points, re-implemented over the C ABI
 * in include/dyncore.h.  Names and signatures are those of michaeldiaz0/Model's headers, so the
 * reference's serial time loop (driver_serial.cpp:178-286) links against this library unchanged.
 *
 * State: the reference keeps every field in host globals (model.h:63-91, microphysics.h:18-25).  Here
 * the fields live in HBM, mirrored from those globals by initialize_pressure_solver() (the first
 * solver call the reference's initialize_serial makes, driver_serial.cpp:58) and copied back by
 * dyn_refapi_download().  Host pointer swaps stay in lock-step with the device buffer rotation.
 *
 * Per-slab flux helpers (compute_fluxes, compute_west, ...) are no-ops: the GPU forms whole stages.
 * Options outside the implemented path (hydrostatic, ice, linear, terrain) report through
 * dyn_last_error() and abort the way the reference aborts on an unsupported option: printf + exit(0).
 */
#ifndef DYNCORE_REFAPI_H
#define DYNCORE_REFAPI_H

#include <cstddef>

struct cell;      /* fluxes.h:34-40 (per-slab scratch; unused here) */
struct vel_cell;  /* fluxes.h:46-62 */

/* ---- pressure.h:7-8 -------------------------------------------------------------------------- */
void solve_pressure(double step);            /* pressure.cpp:142 */
void initialize_pressure_solver();           /* pressure.cpp:99 */

/* ---- advection.h:1-9 ------------------------------------------------------------------------- */
void advect_theta(double step, int il, int ih, int jl, int jh);            /* advection.cpp:349 */
void advect_qv(double step, int il, int ih, int jl, int jh);               /* advection.cpp:577 (linear; unsupported) */
void advect_uv_velocity(double step, int il, int ih, int jl, int jh);      /* advection.cpp:269 (hydrostatic; unsupported) */
void advect_uvw_velocity(double step, int il, int ih, int jl, int jh);     /* advection.cpp:304 */
void advect_microphysics_cell(double step, int il, int ih, int jl, int jh);/* advection.cpp:451 */
void advect_ice_cell(double step, int il, int ih, int jl, int jh);         /* advection.cpp:508 (ice; unsupported) */
void advect_scalar(double, int, int, int, int, double *, double *, double *, struct cell *);
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, double *, struct cell *);
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, struct cell *, struct cell *);

/* ---- fluxes.h:83-102 (slab helpers: no-ops on the GPU path) ------------------------------------ */
void initialize_flux_cells(int ny, int nz);
void initialize_microphysics_cells(int ny, int nz);
void initialize_sign_cells(int nx, int ny, int nz);
void compute_sign_cells(int il, int ih, int jl, int jh);
void compute_fluxes(int i, int jl, int jh);
void compute_west(int i, int jl, int jh);
void compute_west_uv(int i, int jl, int jh);
void compute_fluxes_moisture(int i, int jl, int jh);
void compute_west_moisture(int i, int jl, int jh);

/* ---- microphysics.h:65-93 ---------------------------------------------------------------------- */
void run_microphysics(int il, int ih, int jl, int jh);                  /* microphysics.cpp:97 */
void init_microphysics(int nx, int ny);                                 /* microphysics.cpp:41 */
void microphysics_advance(size_t num_bytes);                            /* microphysics.cpp:657 */
void microphysics_advance_inner(size_t num_bytes);                      /* microphysics.cpp:642 */
void zero_moisture(int il, int ih, int jl, int jh, int size);           /* microphysics.cpp:675 */
void run_kessler_microphysics(int il, int ih, int jl, int jh);          /* kessler.cpp:25 */

/* ---- solver.h:1-7 ------------------------------------------------------------------------------ */
void integrate_non_hydro(double step, int il, int ih, int jl, int jh);  /* solver.cpp:87 */
void integrate_hydro(double step, int itr, int il, int ih, int jl, int jh); /* solver.cpp:35 (unsupported) */
void advance_inner(size_t num_bytes);                                   /* solver.cpp:241 */
void advance_outer(size_t num_bytes);                                   /* solver.cpp:254 */
void switch_array(double **var1, double **var2);                        /* solver.cpp:229 */

/* ---- boundaries.h / damping.h ------------------------------------------------------------------ */
void init_boundaries(int eastLength, int westLength, int northLength, int southLength, int halo); /* boundaries.cpp:44 */
void apply_boundary_condition(int mpi_proc_null);                       /* boundaries.cpp:119 */
void apply_boundary_condition_microphysics(int mpi_proc_null);          /* boundaries.cpp:200 */
void rayleigh_damping(int il, int ih, int jl, int jh, int level);       /* damping.cpp:1148 */

/* ---- the serial RK3 step, driver_serial.cpp:178 (restated over the entry points above) --------- */
void s_integrate_rk3();

/* ---- HBM <-> host-global synchronisation (no reference counterpart) ---------------------------- */
extern "C" {
int dyn_refapi_upload(void);    /* copy every host global field into HBM */
int dyn_refapi_download(void);  /* copy every HBM field back into the host globals */
void *dyn_refapi_context(void); /* the dyn_ctx* the entry points drive */
/* 1 (default): a failed call prints the error and exits with status 0, as the reference does on its
 * fatal paths (kessler.cpp:83-88); 0: the error is kept for dyn_refapi_error() and later calls return */
void dyn_refapi_exit_on_error(int on);
const char *dyn_refapi_error(void);
}

#endif
