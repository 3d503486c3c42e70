/*
 * dyncore.h -- C ABI of the MI355X (gfx950) anelastic dynamical core.
 *
 * Plain pointers and sizes only.  One context owns one model state resident in HBM and one HIP
 * stream; every call is stream-ordered and returns 0 on success or a negative DYN_E_* code, with a
 * message from dyn_last_error().  The C++ drop-in layer (include/dyncore_refapi.h) keeps the
 * reference's own entry points and forwards to these calls.
 *
 * Grid conventions are the reference's (model.h:46-58): arrays are [i][j][k] with k fastest,
 * NX includes three periodic halo columns on each side, NZ includes the ghost levels 0 and NZ-1.
 * Bounds arguments il, ih, jl, jh are half-open ranges in array coordinates, exactly as the
 * reference functions take them.
 */
#ifndef DYNCORE_H
#define DYNCORE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- identifiers --------------------------------------------------------------------------- */
enum dyn_field {
    DYN_U = 0, DYN_V, DYN_W, DYN_TH, DYN_QV, DYN_QC, DYN_QR, /* prognostics: 3 buffers each */
    DYN_PI,                                                  /* perturbation pressure (pis) */
    DYN_UBAR, DYN_VBAR, DYN_WBAR, DYN_THBAR, DYN_QBAR, DYN_PBAR, /* basic state (m_ubar ...) */
    DYN_VT,                                                  /* rain fall speed (vts) */
    DYN_ACCRAIN,                                             /* accumulated rain [NX][NY] */
    DYN_FIELD_COUNT
};
/* the three rotating buffers of a prognostic: ums / us / ups (model.h:63-65) */
enum dyn_role { DYN_ROLE_M = 0, DYN_ROLE_CUR = 1, DYN_ROLE_P = 2 };
enum dyn_column {
    DYN_COL_ZU = 0, DYN_COL_ZW, DYN_COL_ZSU, DYN_COL_ZSW, DYN_COL_MU, DYN_COL_MW,
    DYN_COL_RHOU, DYN_COL_RHOW, DYN_COL_ONE_D_RHOU, DYN_COL_ONE_D_RHOW,
    DYN_COL_TB, DYN_COL_TBV, DYN_COL_TBW, DYN_COL_QB, DYN_COL_PIB,
    DYN_COL_F, /* Coriolis parameter f[j], length NY */
    DYN_COLUMN_COUNT
};
enum dyn_precision { DYN_FP64 = 0, DYN_FP32 = 1 };

/* error codes */
#define DYN_OK 0
#define DYN_E_ARG (-1)         /* invalid argument (bounds, ids, sizes) */
#define DYN_E_UNSUPPORTED (-2) /* option outside the implemented path (e.g. non-periodic east-west) */
#define DYN_E_HIP (-3)         /* HIP runtime failure */
#define DYN_E_NUMERIC (-4)     /* the reference's fatal Kessler guard E>100 (kessler.cpp:83-88) or NaN */
#define DYN_E_STATE (-5)       /* call out of order (e.g. solve before the columns are set) */

typedef struct dyn_grid {
    int nx, ny, nz;          /* NX (incl. 6 halo columns), NY, NZ (incl. 2 ghost levels) */
    double dx, dz, dt;       /* dy = dx, as process_input.cpp:175-176 sets it */
    int microphysics;        /* MICROPHYSICS_OPTION: 0 dry, 1 Kessler (2, ice, is unsupported) */
    int periodic_ew;         /* PERIODIC_BOUNDARIES: must be 1 (serial periodic path) */
    int rayleigh_level;      /* raydampheight (initializer.cpp:185); >= nz disables damping */
    int sponge_north, sponge_south; /* jnbuffer, jsbuffer (boundaries.cpp:23-26) */
    int precision;           /* DYN_FP64 (plain form) or DYN_FP32 (fp32 storage, fp64 vertical solve) */
} dyn_grid;

typedef struct dyn_ctx dyn_ctx;

const char *dyn_last_error(void);
int dyn_version(void);

/* lifetime: initialize_pressure_solver + init_boundaries + setup_memory_allocation equivalents */
int dyn_create(const dyn_grid *grid, dyn_ctx **out);
int dyn_destroy(dyn_ctx *ctx);

/* host <-> HBM transfers (double on the host side for both precisions) */
int dyn_set_column(dyn_ctx *ctx, int column, const double *values, size_t count);
int dyn_set_field(dyn_ctx *ctx, int field, int role, const double *values, size_t count);
int dyn_get_field(dyn_ctx *ctx, int field, int role, double *values, size_t count);
int dyn_synchronize(dyn_ctx *ctx);
/* Failure detection (no counterpart in the reference, which has only the Kessler guard of
 * kessler.cpp:83-88, reported by dyn_synchronize): scans the current u, v, w, theta (qv, qc, qr) and pi
 * for NaN or Inf; returns DYN_E_NUMERIC naming the fields, with bit DYN_<field> of *bad_fields set for
 * each (bad_fields may be NULL).  One read of each field; stream-ordered, then synchronises. */
int dyn_check_finite(dyn_ctx *ctx, int *bad_fields);
/* dyn_get_field without any device allocation.  No counterpart in the reference: it is for callers that
 * cannot allocate device memory (tight serving or output loops, harnesses that forbid hipMalloc).  An
 * FP32 field is widened to double in a file-scope device array of 8,388,608 doubles (64 MiB, one per
 * GPU, held until the process ends); fields of more values return DYN_E_ARG (use dyn_get_field).  Every
 * call on one GPU shares that array, so calls that overlap on one GPU (from several host threads)
 * race.  FP64 fields copy straight to the host.  Same values as dyn_get_field, bit for bit. */
int dyn_get_field_noalloc(dyn_ctx *ctx, int field, int role, double *values, size_t count);

/* ---- reference operators, one call each -------------------------------------------------- */
/* advect_uvw_velocity(step,il,ih,jl,jh), advection.cpp:304 */
int dyn_advect_uvw(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* advect_theta(step,il,ih,jl,jh), advection.cpp:349 */
int dyn_advect_theta(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* advect_microphysics_cell(step,il,ih,jl,jh), advection.cpp:451 (includes the fall speed) */
int dyn_advect_moisture(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* zero_moisture(il,ih,jl,jh,size), microphysics.cpp:675 */
int dyn_zero_moisture(dyn_ctx *ctx, int il, int ih, int jl, int jh);
/* solve_pressure(step), pressure.cpp:142 (serial periodic: uvw halos, poisson_fft3d, pi halos) */
int dyn_solve_pressure(dyn_ctx *ctx, double step);
/* pressure-gradient force of integrate_non_hydro, solver.cpp:123-135, over [il,ih) x [jl,jh) */
int dyn_pgf(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* apply_boundary_condition(0), boundaries.cpp:119 / apply_boundary_condition_microphysics(0), :200 */
int dyn_apply_bc(dyn_ctx *ctx);
int dyn_apply_bc_micro(dyn_ctx *ctx);
/* run_microphysics(il,ih,jl,jh) -> run_kessler_microphysics + precip_rate, kessler.cpp:25 */
int dyn_kessler(dyn_ctx *ctx, int il, int ih, int jl, int jh);
/* rayleigh_damping(il,ih,jl,jh,level), damping.cpp:1148 */
int dyn_rayleigh(dyn_ctx *ctx, int il, int ih, int jl, int jh, int level);
/* advance_inner / advance_outer, solver.cpp:241-263; microphysics_advance(_inner), microphysics.cpp:642-669 */
int dyn_advance_inner(dyn_ctx *ctx);
int dyn_advance_outer(dyn_ctx *ctx);
int dyn_micro_advance_inner(dyn_ctx *ctx);
int dyn_micro_advance_outer(dyn_ctx *ctx);

/* ---- turbulence parameterization (USE_TURBULENT_STRESS, turbulence.cpp) ------------------------ */
/* Switch the Smagorinsky-Lilly closure on or off (initialize_landsea + init_kmix, turbulence.cpp:85-147,
 * 370-375, with the vertical mixing length from the zsu column).  landsea holds NX*NY values (< 0.5:
 * land, fixed drag 3e-3; otherwise ocean, 1.1e-3 + 4e-5 |V|) or is NULL for all ocean.  With it on, a
 * step computes the coefficients and tendencies at its start and adds them in every stage; steps take
 * the per-operator sequence (the fused stage kernels carry no turbulence), and the adjoint and
 * ensembles refuse the context.  Drops the captured step graphs. */
int dyn_set_turbulence(dyn_ctx *ctx, int on, const double *landsea, size_t count);
/* calculate_diff_tend(il,ih,jl,jh), turbulence.cpp:153-226 (the serial driver calls it over
 * [1,NX-1) x [1,NY-1), driver_serial.cpp:194): mixing coefficients from the M buffers, then the
 * friction and scalar diffusion tendencies */
int dyn_calculate_diff_tend(dyn_ctx *ctx, int il, int ih, int jl, int jh);
/* apply_velocity_diffusion(step,il,ih,jl,jh), turbulence.cpp:233-258 (solver.cpp:104) */
int dyn_apply_velocity_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* apply_scalar_diffusion(step,il,ih,jl,jh), turbulence.cpp:263-273 (driver_serial.cpp:218) */
int dyn_apply_scalar_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* apply_moisture_diffusion(step,il,ih,jl,jh), turbulence.cpp:278-292 (driver_serial.cpp:231) */
int dyn_apply_moisture_diffusion(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh);
/* read one turbulence array as FP64 (NX*NY*NZ values): 0 Kmix, 1 KHmix, 2-4 u/v/w friction, 5-8
 * theta/qv/qc/qr diffusion tendencies (turbulence.cpp:32-47), 9 Kessler's saturation flag (0 or 1,
 * isSaturated, kessler.cpp:129-150) */
int dyn_get_turbulence(dyn_ctx *ctx, int which, double *values, size_t count);

/* ---- whole steps ------------------------------------------------------------------------- */
/* s_integrate_rk3 (driver_serial.cpp:178-286) as the sequence of per-operator calls above */
int dyn_step_reference(dyn_ctx *ctx, int nsteps);
/* nsteps steps replayed from captured hipGraphs: when NZ <= 64 or 66 <= NZ <= 128 the stage advection
 * runs as the fused momentum and scalar kernels with fused boundary epilogues (same values as the
 * per-operator kernels, bit for bit, in either precision); otherwise the per-operator sequence is
 * replayed.  Between the steps of one call the M and current roles share a buffer (no M <- P copy), so
 * one call of n steps is cheaper than n calls of one; every call ends with the three buffers apart and
 * the values n single steps leave. */
int dyn_step(dyn_ctx *ctx, int nsteps);
/* the fused-stage step launched eagerly, with the same sharing between the steps of the call (either
 * precision; needs NZ <= 64 or 66 <= NZ <= 128, and fields below 4 GiB) */
int dyn_step_fused(dyn_ctx *ctx, int nsteps);

/* ---- measurement ------------------------------------------------------------------------- */
/* HBM bytes the step's algorithm must move per gridpoint (the field-pass count of the operator set) */
double dyn_algorithmic_bytes_per_point(const dyn_ctx *ctx);
/* per-kernel HIP-event timing of eager launches (dyn_step_reference and the per-operator calls);
 * dyn_timing(ctx, 1) clears the records and starts timing, dyn_timing_report writes one line per kernel:
 * "<name> <launches> <total_ms> <gridpoints processed>" */
int dyn_timing(dyn_ctx *ctx, int on);
int dyn_timing_report(dyn_ctx *ctx, char *buf, size_t len);
/* the HIP stream the context launches on (a hipStream_t) */
void *dyn_stream(dyn_ctx *ctx);

/* ---- launch-shape parameters ------------------------------------------------------------- */
/* Knobs of the fused kernels (no counterpart in the reference: they set how the work maps onto the
 * GPU, never what is computed).  0 restores the built-in default; a value the kernels do not support
 * returns DYN_E_ARG.  Setting a knob drops the captured step graph. */
enum {
    DYN_TUNE_MOMENTUM_TJ = 0, /* j-rows per block of the momentum stage kernel: 2, 4, 8 or 16 (16, like 8,
                                 only with one wavefront per column) */
    DYN_TUNE_MOMENTUM_TI = 1, /* columns one momentum block marches along i: 2 .. 512 */
    DYN_TUNE_SCALAR_TJ = 2,   /* j-rows per block of the scalar (theta, moisture) stage kernel: 2, 4, 8 or 12
                                 (8 and 12 only with one wavefront per column).  Default 8, or 4 for
                                 small FP64 grids; FP32 with the stage on one stream (DYN_TUNE_STAGE_STREAMS
                                 set to 1) takes 12 */
    DYN_TUNE_SCALAR_TI = 3,   /* columns one scalar block marches along i: 2 .. 512.  With 12 rows the
                                 default fills the device with one wave of blocks (at least 16) */
    DYN_TUNE_PRESSURE_JPASS = 4, /* pressure j pass: 1 one LDS-resident kernel per kx (when the slab fits), 2 three
                                    kernels, 3 two LDS-resident kernels per kx on half slabs (even NY, when they fit) */
    DYN_TUNE_STAGE_STREAMS = 5,  /* fused step: 1 runs the whole stage on the context's stream; 2 runs a
                                    stage's scalar half on a second stream beside the pressure solve, forked
                                    after the momentum kernel; 3 the same, forked after the solve's first
                                    pass; 4 forked after the momentum kernel, its launches queued after the
                                    solve's first pass.  Default: 3 where the j pass runs on half slabs,
                                    else 2 (FP32) or 4 (FP64) */
    DYN_TUNE_COUNT = 6
};
int dyn_tune_set(dyn_ctx *ctx, int param, int value);
int dyn_tune_get(dyn_ctx *ctx, int param);

/* ---- output: asynchronous staging and the reference's NetCDF file ------------------------------- */
/* The reference writes u, v, w, pi, theta (and qv, qc, qr, rainfall when moist) every outfilefreq steps
 * from its host arrays (driver_serial.cpp:317-345 -> output_meteorological_fields_to_file,
 * files.cpp:673-700).  dyn_output_stage narrows the current values of those fields to float in a device
 * buffer (stream-ordered after the steps queued before it) and copies them to pinned host memory on a
 * stream of its own, so the steps queued after it overlap the copy; the step loop never waits on the
 * host.  dyn_output_ready polls (1 landed, 0 not yet); dyn_output_get (one field) and
 * dyn_netcdf_write_staged (one record of the file) wait for the copy.  The device and host buffers are
 * allocated at the first dyn_output_stage and kept by the context. */
int dyn_output_stage(dyn_ctx *ctx);
int dyn_output_ready(dyn_ctx *ctx);
int dyn_output_get(dyn_ctx *ctx, int field, float *values, size_t count);

/* The reference's output file, written by this library's own netCDF classic-format (64-bit offset)
 * writer: create_outfile's dimensions, global attributes and NC_FLOAT variables (files.cpp:51-282).
 * The attributes the serial periodic path fixes (hydrostatic, islinear, stretched_grid ...) are written
 * with those values; the rest come from dyn_nc_meta. */
#define DYN_NC_MODEL_FILES_NOT_WRITTEN 0 /* file.h:15-17: values of last_file_successfully_written */
#define DYN_NC_MODEL_FILES_WRITTEN 1
#define DYN_NC_ALL_FILES_WRITTEN 2
typedef struct dyn_nc_meta {
    double dx, dz, dt;                     /* "dx" and "dy" (dy = dx), "dz", "dt" */
    double latoffset, lonoffset, water_temperature, height_lowest_level;
    int use_turbulent_stress, basic_state_option, perturbation_option;
    const char *reanalysis_input_data, *perturbation_input_data, *basic_state_input_data; /* NULL: "" */
} dyn_nc_meta;
/* create_outfile(path, true, true, true) for an NX x NY x NZ grid (moisture variables when microphysics
 * != 0), variables zero-filled, no records, last_file_successfully_written = DYN_NC_ALL_FILES_WRITTEN */
int dyn_netcdf_create(const char *path, int nx, int ny, int nz, int microphysics, const dyn_nc_meta *meta);
/* outfile_init (files.cpp:1262-1285) for a context: dyn_netcdf_create plus the basic state
 * (write_all_bvars), tb, pib, qb, zu = zsu (write_all_mbvars) and lat[NY], lon[NX] (NULL: zeros) */
int dyn_netcdf_init(dyn_ctx *ctx, const char *path, const dyn_nc_meta *meta, const double *lat, const double *lon);
/* one record of the staged fields and time = hours (output_meteorological_fields_to_file and
 * write_time_to_file, files.cpp:628-700); extends the record count */
int dyn_netcdf_write_staged(dyn_ctx *ctx, const char *path, int record, double hours);
/* nc_put_vara_double / nc_get_vara_double of a whole variable (one record of a record variable), as
 * doubles; count must be the variable's value count per record.  Reads any classic-format file. */
int dyn_netcdf_put(const char *path, const char *var, int record, const double *values, size_t count);
int dyn_netcdf_read(const char *path, const char *var, int record, double *values, size_t count);
int dyn_netcdf_records(const char *path); /* the record count (>= 0) or a DYN_E_* code */
/* file_output_status / get_file_output_status (files.cpp:1205-1254): the restart attribute
 * last_file_successfully_written; get returns -1 in *status when the attribute is absent */
int dyn_netcdf_set_status(const char *path, int status);
int dyn_netcdf_get_status(const char *path, int *status);

/* ---- discrete adjoint of the dry step ------------------------------------------------------ */
/* No counterpart in the reference (it has no adjoint).  dyn_adjoint_record runs nsteps forward steps
 * from the current state (leaving the context at the final state) and keeps a snapshot of the u, v, w,
 * theta buffers (M, current, forecast) every `every` steps.  dyn_adjoint_backward takes the gradient
 * of a scalar J with respect to the final current u, v, w, theta (host arrays of NX*NY*NZ doubles, NULL
 * meaning zero) and returns dJ/d(initial field) for each (NULL to skip), where an initial field is the
 * value set into all three buffers of that prognostic.  It recomputes each segment from its snapshot
 * and runs the transposed kernels backwards; upwind signs are those of the recorded trajectory.  The
 * context is left at the final state again.  Dry contexts only (microphysics 0); the trajectory memory
 * is 12 fields per snapshot plus 12 fields per step of one segment.  The record pass keeps the stage
 * states of the last segment, so that segment is never recomputed: with every >= nsteps the reverse
 * sweep recomputes nothing.  every = 0 picks the longest segment whose memory fits in half of the free
 * device memory; dyn_adjoint_interval returns the segment length in use. */
int dyn_adjoint_record(dyn_ctx *ctx, int nsteps, int every);
int dyn_adjoint_interval(dyn_ctx *ctx);
int dyn_adjoint_backward(dyn_ctx *ctx, const double *seed_u, const double *seed_v, const double *seed_w, const double *seed_th,
                         double *grad_u, double *grad_v, double *grad_w, double *grad_th);
int dyn_adjoint_release(dyn_ctx *ctx);

/* ---- ensembles ----------------------------------------------------------------------------- */
/* No counterpart in the reference (its ensemble driver runs members one after another on the CPU).
 * An ensemble advances existing contexts of possibly different NX, NY (same precision and NZ, dry,
 * NZ <= 64) together: every pass of the step is one launch over all members' work, and
 * dyn_ensemble_step replays a captured hipGraph.  Members must outlive the ensemble; call
 * dyn_ensemble_synchronize before reading or writing a member's fields.  Per member the results are
 * bit for bit those of dyn_step on that member alone. */
typedef struct dyn_ensemble dyn_ensemble;
int dyn_ensemble_create(dyn_ctx *const *members, int n, dyn_ensemble **out);
int dyn_ensemble_destroy(dyn_ensemble *ens);
int dyn_ensemble_synchronize(dyn_ensemble *ens);
int dyn_ensemble_step(dyn_ensemble *ens, int nsteps);
int dyn_ensemble_step_eager(dyn_ensemble *ens, int nsteps);
/* launch shape {momentum rows per block, momentum march length, scalar rows, scalar march length} */
int dyn_ensemble_set_shape(dyn_ensemble *ens, const int *shape);
/* times every shape of the two stage kernels on the members' grids and keeps the fastest (written to
 * shape_out[4] when not NULL); the model state is unchanged */
int dyn_ensemble_autotune(dyn_ensemble *ens, int *shape_out);
/* splits the members into ngroups groups of about equal point counts, each advanced by the same batched
 * launches on a stream of its own (forked from and joined back into the first member's stream, one
 * branch per group in the captured graph), so that different passes of different groups overlap.
 * Results are unchanged.  dyn_ensemble_groups returns the count in force (min(4, members) after create). */
int dyn_ensemble_set_groups(dyn_ensemble *ens, int ngroups);
int dyn_ensemble_groups(dyn_ensemble *ens);

#ifdef __cplusplus
}
#endif
#endif
