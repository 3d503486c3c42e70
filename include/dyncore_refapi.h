/*
 * dyncore_refapi.h -- the reference's own dynamical-core entry points, re-implemented over the C ABI
 * in include/dyncore.h.  Names and signatures are those of michaeldiaz0/Model's headers, so the
 * reference's serial time loop (driver_serial.cpp:178-286) calls them unchanged.
 *
 * Linking: the reference keeps every one of its objects.  In the objects that define the functions
 * declared below (pressure.o, advection.o, fluxes.o, kessler.o, microphysics.o, boundaries.o,
 * damping.o, solver.o, turbulence.o) these symbols are made local with objcopy --localize-symbols, using the list
 * the build writes next to the library (libdyncore_refapi.replaced: the C++ symbols this library
 * defines); every call from another object then binds to this library.  Everything else in those
 * objects stays as it is (init_boundaries, init_damping, init_microphysics, precip_rate, the fall-speed
 * functions, init_kmix, initialize_landsea, the output code ...).
 *
 * State: the reference keeps every field in host globals (model.h:63-91, microphysics.h:18-25), which
 * its own initializer.cpp defines; this library only references them.  The fields live in HBM,
 * mirrored from those globals by initialize_pressure_solver() (the first solver call the reference's
 * initialize_serial makes, driver_serial.cpp:58) and copied back by dyn_refapi_download().  Host
 * pointer swaps stay in lock-step with the device buffer rotation.
 *
 * Options outside the implemented path (hydrostatic, ice, linear, tracers, semi-Lagrangian rain)
 * report through dyn_refapi_error() and abort the way the reference aborts on an
 * unsupported option: printf + exit(0).
 */
#ifndef DYNCORE_REFAPI_H
#define DYNCORE_REFAPI_H

#include <cstddef>

struct cell; /* fluxes.h:34-40 (per-slab scratch; unused here) */

/* ---- pressure.h:7-8 -------------------------------------------------------------------------- */
void solve_pressure(double step);            /* pressure.cpp:142 */
void initialize_pressure_solver();           /* pressure.cpp:99 (also creates the HBM state) */

/* ---- advection.h:1-9 ------------------------------------------------------------------------- */
void advect_theta(double step, int il, int ih, int jl, int jh);            /* advection.cpp:349 */
void advect_uvw_velocity(double step, int il, int ih, int jl, int jh);     /* advection.cpp:304 */
void advect_microphysics_cell(double step, int il, int ih, int jl, int jh);/* advection.cpp:451 */
void advect_qv(double step, int il, int ih, int jl, int jh);               /* advection.cpp:577 (linear; unsupported) */
void advect_uv_velocity(double step, int il, int ih, int jl, int jh);      /* advection.cpp:269 (hydrostatic; unsupported) */
void advect_ice_cell(double step, int il, int ih, int jl, int jh);         /* advection.cpp:508 (ice; unsupported) */
/* tracer advection, advection.cpp:383 and :417, and the overload advection.h:8 declares (unsupported) */
void advect_scalar(double, int, int, int, int, double *, double *, double *, struct cell *);
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, double *, struct cell *);
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, struct cell *, struct cell *);

/* ---- fluxes.h:86 ------------------------------------------------------------------------------ */
/* compute_sign_cells, fluxes.cpp:316: a no-op; every stage kernel evaluates the signs it needs.  The
 * per-slab helpers (compute_fluxes, compute_west, interpolate_*) are not replaced: only the replaced
 * advection functions call them. */
void compute_sign_cells(int il, int ih, int jl, int jh);

/* ---- microphysics.h:65-93, kessler.cpp ---------------------------------------------------------- */
void run_microphysics(int il, int ih, int jl, int jh);                  /* microphysics.cpp:97 */
void microphysics_advance(size_t num_bytes);                            /* microphysics.cpp:657 */
void microphysics_advance_inner(size_t num_bytes);                      /* microphysics.cpp:642 */
void zero_moisture(int il, int ih, int jl, int jh, int size);           /* microphysics.cpp:675 */
void run_kessler_microphysics(int il, int ih, int jl, int jh);          /* kessler.cpp:25 (+ precip_rate) */

/* ---- solver.h:1-7 ------------------------------------------------------------------------------ */
void integrate_non_hydro(double step, int il, int ih, int jl, int jh);  /* solver.cpp:87 */
void integrate_hydro(double step, int itr, int il, int ih, int jl, int jh); /* solver.cpp:35 (unsupported) */
void advance_inner(size_t num_bytes);                                   /* solver.cpp:241 */
void advance_outer(size_t num_bytes);                                   /* solver.cpp:254 */

/* ---- turbulence.h:8-11 (USE_TURBULENT_STRESS; the closure's state is created by
 * initialize_pressure_solver from the land-sea mask initialize_landsea left) ---------------------- */
void calculate_diff_tend(int il, int ih, int jl, int jh);                  /* turbulence.cpp:158 */
void apply_velocity_diffusion(double step, int il, int ih, int jl, int jh); /* turbulence.cpp:233 */
void apply_scalar_diffusion(double step, int il, int ih, int jl, int jh);   /* turbulence.cpp:263 */
void apply_moisture_diffusion(double step, int il, int ih, int jl, int jh); /* turbulence.cpp:278 */

/* ---- boundaries.h / damping.h ------------------------------------------------------------------ */
void apply_boundary_condition(int mpi_proc_null);                       /* boundaries.cpp:119 */
void apply_boundary_condition_microphysics(int mpi_proc_null);          /* boundaries.cpp:200 */
void rayleigh_damping(int il, int ih, int jl, int jh, int level);       /* damping.cpp:1148 */

/* ---- HBM <-> host-global synchronisation (no reference counterpart) ---------------------------- */
extern "C" {
int dyn_refapi_upload(void);    /* copy every host global field into HBM */
int dyn_refapi_download(void);  /* copy every HBM field back into the host globals */
void *dyn_refapi_context(void); /* the dyn_ctx* the entry points drive */
/* 1 (default): a failed call prints the error and exits with status 0, as the reference does on its
 * fatal paths (kessler.cpp:83-88); 0: the error is kept for dyn_refapi_error() and later calls return */
void dyn_refapi_exit_on_error(int on);
const char *dyn_refapi_error(void);
void dyn_refapi_reset(void);    /* destroy the context; the next initialize_pressure_solver makes a new one */
}

#endif
