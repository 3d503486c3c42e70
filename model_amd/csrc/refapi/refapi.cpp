// The reference's dynamical-core entry points over the C ABI of libdyncore.
//
// Each function keeps the name, signature and argument meaning of the reference function it replaces
// (cited per function) and forwards to the matching dyn_* call.  Device buffers rotate together with
// the host pointer swaps the reference performs (switch_array, solver.cpp:229-235), so host globals and
// HBM stay paired by role.  The host globals are the reference's own (initializer.cpp defines them);
// this library only references them.
#include <cstdio>
#include <cstdlib>
#include <string>
#include <utility>

#include "../../../include/dyncore.h"
#include "../../../include/dyncore_refapi.h"
#include "refapi_globals.h"

namespace {

dyn_ctx *g_ctx = nullptr;
bool g_exit_on_error = true;
std::string g_error;

void fatal(const std::string &what) {
    g_error = what;
    if (g_exit_on_error) {
        printf("%s\n", what.c_str());
        fflush(stdout);
        exit(0);
    }
}

bool ok() { return g_ctx != nullptr && g_error.empty(); }

void check(int rc, const char *call) {
    if (rc != DYN_OK) fatal(std::string(call) + ": " + dyn_last_error());
}

#define CALL(expr)                  \
    do {                            \
        if (!ok()) return;          \
        check((expr), #expr);       \
    } while (0)

void swap_host(double **a, double **b) {
    double *t = *a;
    *a = *b;
    *b = t;
}

size_t field_size() { return (size_t)NX * NY * NZ; }

// host pointer of a field in a role (the pairing the device buffers follow)
double *host_field(int field, int role) {
    switch (field) {
    case DYN_U: return role == DYN_ROLE_M ? ums : role == DYN_ROLE_CUR ? us : ups;
    case DYN_V: return role == DYN_ROLE_M ? vms : role == DYN_ROLE_CUR ? vs : vps;
    case DYN_W: return role == DYN_ROLE_M ? wms : role == DYN_ROLE_CUR ? ws : wps;
    case DYN_TH: return role == DYN_ROLE_M ? thms : role == DYN_ROLE_CUR ? ths : thps;
    case DYN_QV: return role == DYN_ROLE_M ? qvms : role == DYN_ROLE_CUR ? qvs : qvps;
    case DYN_QC: return role == DYN_ROLE_M ? qcms : role == DYN_ROLE_CUR ? qcs : qcps;
    case DYN_QR: return role == DYN_ROLE_M ? qrms : role == DYN_ROLE_CUR ? qrs : qrps;
    case DYN_PI: return pis;
    case DYN_UBAR: return m_ubar;
    case DYN_VBAR: return m_vbar;
    case DYN_WBAR: return m_wbar;
    case DYN_THBAR: return m_thbar;
    case DYN_QBAR: return m_qbar;
    case DYN_PBAR: return m_pbar;
    case DYN_VT: return vts;
    case DYN_ACCRAIN: return accRain;
    }
    return nullptr;
}

int transfer_all(bool up) {
    int nprog = USE_MICROPHYSICS ? 7 : 4;
    for (int f = 0; f < DYN_FIELD_COUNT; f++) {
        bool prog = f <= DYN_QR;
        if (prog && f >= nprog) continue;
        if ((f == DYN_VT || f == DYN_ACCRAIN || f == DYN_QBAR) && !USE_MICROPHYSICS) continue;
        size_t n = f == DYN_ACCRAIN ? (size_t)NX * NY : field_size();
        for (int r = 0; r < (prog ? 3 : 1); r++) {
            double *h = host_field(f, r);
            if (!h) continue;
            int rc = up ? dyn_set_field(g_ctx, f, r, h, n) : dyn_get_field(g_ctx, f, r, h, n);
            if (rc) return rc;
        }
    }
    return DYN_OK;
}

void unsupported(const char *name) { fatal(std::string(name) + ": option outside the implemented dynamical-core path"); }

}  // namespace

extern "C" {
int dyn_refapi_upload(void) { return g_ctx ? transfer_all(true) : DYN_E_STATE; }
int dyn_refapi_download(void) {
    if (!g_ctx) return DYN_E_STATE;
    int rc = dyn_synchronize(g_ctx);
    return rc ? rc : transfer_all(false);
}
void *dyn_refapi_context(void) { return g_ctx; }
void dyn_refapi_exit_on_error(int on) { g_exit_on_error = on != 0; }
const char *dyn_refapi_error(void) { return g_error.c_str(); }
void dyn_refapi_reset(void) {
    if (g_ctx) dyn_destroy(g_ctx);
    g_ctx = nullptr;
    g_error.clear();
}
}

// ---- pressure.h -----------------------------------------------------------------------------------
// initialize_pressure_solver, pressure.cpp:99-134: here it also creates the HBM state from the globals
void initialize_pressure_solver() {
    if (g_ctx) dyn_refapi_reset();
    g_error.clear();
    dyn_grid g{};
    g.nx = NX;
    g.ny = NY;
    g.nz = NZ;
    g.dx = dx;
    g.dz = dz;
    g.dt = dt;
    g.microphysics = MICROPHYSICS_OPTION;
    g.periodic_ew = PERIODIC_BOUNDARIES;
    g.rayleigh_level = raydampheight;
    g.sponge_north = jnbuffer;
    g.sponge_south = jsbuffer;
    g.precision = DYN_FP64;
    dyn_ctx *c = nullptr;
    int rc = dyn_create(&g, &c);
    if (rc) {
        fatal(std::string("initialize_pressure_solver: ") + dyn_last_error());
        return;
    }
    g_ctx = c;
    double *cols[DYN_COLUMN_COUNT] = {zu, zw, zsu, zsw, mu, mw, rhou, rhow, one_d_rhou, one_d_rhow, tb, tbv, tbw, qb, pib, f};
    for (int k = 0; k < DYN_COLUMN_COUNT; k++) CALL(dyn_set_column(g_ctx, k, cols[k], k == DYN_COL_F ? (size_t)NY : (size_t)NZ));
    CALL(dyn_refapi_upload());
    // the reference's initialize_serial runs init_kmix and initialize_landsea before this call
    // (driver_serial.cpp:52-58): the closure's device state starts here, from the mask those left
    if (USE_TURBULENT_STRESS) CALL(dyn_set_turbulence(g_ctx, 1, landsea, landsea ? (size_t)NX * NY : 0));
}

void solve_pressure(double step) { CALL(dyn_solve_pressure(g_ctx, step)); }  // pressure.cpp:142

// ---- advection.h ----------------------------------------------------------------------------------
void advect_uvw_velocity(double step, int il, int ih, int jl, int jh) { CALL(dyn_advect_uvw(g_ctx, step, il, ih, jl, jh)); }
void advect_theta(double step, int il, int ih, int jl, int jh) { CALL(dyn_advect_theta(g_ctx, step, il, ih, jl, jh)); }
void advect_microphysics_cell(double step, int il, int ih, int jl, int jh) {
    if (RAIN_FALLOUT != 1) return unsupported("advect_microphysics_cell (semi-Lagrangian rain fallout)");
    CALL(dyn_advect_moisture(g_ctx, step, il, ih, jl, jh));
}
void advect_qv(double, int, int, int, int) { unsupported("advect_qv"); }
void advect_uv_velocity(double, int, int, int, int) { unsupported("advect_uv_velocity"); }
void advect_ice_cell(double, int, int, int, int) { unsupported("advect_ice_cell"); }
void advect_scalar(double, int, int, int, int, double *, double *, double *, struct cell *) { unsupported("advect_scalar"); }
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, double *, struct cell *) {
    unsupported("advect_scalar");
}
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, struct cell *, struct cell *) {
    unsupported("advect_scalar");
}

// ---- fluxes.h: sign cells are evaluated inside the stage kernels --------------------------------------
void compute_sign_cells(int, int, int, int) {}  // fluxes.cpp:316

// ---- microphysics.h --------------------------------------------------------------------------------
// run_microphysics, microphysics.cpp:97-132
void run_microphysics(int il, int ih, int jl, int jh) {
    switch (MICROPHYSICS_OPTION) {
    case 0: break;
    case 1: run_kessler_microphysics(il, ih, jl, jh); break;
    case 2: unsupported("run_rutledge_microphysics"); break;
    default: printf("Error! Microphysics option %d not supported", MICROPHYSICS_OPTION); break;
    }
}

void run_kessler_microphysics(int il, int ih, int jl, int jh) {  // kessler.cpp:25-213 (+ precip_rate)
    if (RAIN_FALLOUT != 1) return unsupported("run_kessler_microphysics (semi-Lagrangian rain fallout)");
    CALL(dyn_kessler(g_ctx, il, ih, jl, jh));
}

void zero_moisture(int il, int ih, int jl, int jh, int size) {  // microphysics.cpp:675-700
    if ((size_t)size != field_size()) return fatal("zero_moisture: size must be NX*NY*NZ");
    CALL(dyn_zero_moisture(g_ctx, il, ih, jl, jh));
}

void microphysics_advance_inner(size_t) {  // microphysics.cpp:642-652
    if (!ok()) return;
    swap_host(&qvs, &qvps);
    swap_host(&qcs, &qcps);
    swap_host(&qrs, &qrps);
    check(dyn_micro_advance_inner(g_ctx), "dyn_micro_advance_inner");
}

void microphysics_advance(size_t) {  // microphysics.cpp:657-669
    if (!ok()) return;
    swap_host(&qvs, &qvps);
    swap_host(&qcs, &qcps);
    swap_host(&qrs, &qrps);
    check(dyn_micro_advance_outer(g_ctx), "dyn_micro_advance_outer");
}

// ---- solver.h --------------------------------------------------------------------------------------
// integrate_non_hydro, solver.cpp:87-138 (serial: buffer = 1)
void integrate_non_hydro(double step, int il, int ih, int jl, int jh) {
    const int buffer = 1;
    advect_uvw_velocity(step, il + 3, ih - 3, jl + 3, jh - 3);
    if (USE_TURBULENT_STRESS) apply_velocity_diffusion(step, il + 3, ih - 3, jl + 3, jh - 3);  // solver.cpp:103
    solve_pressure(step);
    CALL(dyn_pgf(g_ctx, step, il + buffer, ih - buffer, jl + buffer, jh - buffer));
}

void integrate_hydro(double, int, int, int, int, int) { unsupported("integrate_hydro"); }

void advance_inner(size_t) {  // solver.cpp:241-248
    if (!ok()) return;
    swap_host(&us, &ups);
    swap_host(&vs, &vps);
    swap_host(&ths, &thps);
    swap_host(&ws, &wps);
    check(dyn_advance_inner(g_ctx), "dyn_advance_inner");
}

void advance_outer(size_t) {  // solver.cpp:254-263: memcpy M<-P on the device, then the swap
    if (!ok()) return;
    swap_host(&us, &ups);
    swap_host(&vs, &vps);
    swap_host(&ths, &thps);
    swap_host(&ws, &wps);
    check(dyn_advance_outer(g_ctx), "dyn_advance_outer");
}

// ---- turbulence.h ------------------------------------------------------------------------------------
// calculate_diff_tend, turbulence.cpp:158-190: coefficients and tendencies from the M buffers in HBM
void calculate_diff_tend(int il, int ih, int jl, int jh) { CALL(dyn_calculate_diff_tend(g_ctx, il, ih, jl, jh)); }
// apply_velocity_diffusion / apply_scalar_diffusion / apply_moisture_diffusion, turbulence.cpp:233-292
void apply_velocity_diffusion(double step, int il, int ih, int jl, int jh) {
    CALL(dyn_apply_velocity_diffusion(g_ctx, step, il, ih, jl, jh));
}
void apply_scalar_diffusion(double step, int il, int ih, int jl, int jh) { CALL(dyn_apply_scalar_diffusion(g_ctx, step, il, ih, jl, jh)); }
void apply_moisture_diffusion(double step, int il, int ih, int jl, int jh) {
    CALL(dyn_apply_moisture_diffusion(g_ctx, step, il, ih, jl, jh));
}

// ---- boundaries.h / damping.h ------------------------------------------------------------------------
void apply_boundary_condition(int) { CALL(dyn_apply_bc(g_ctx)); }
void apply_boundary_condition_microphysics(int) { CALL(dyn_apply_bc_micro(g_ctx)); }
void rayleigh_damping(int il, int ih, int jl, int jh, int level) { CALL(dyn_rayleigh(g_ctx, il, ih, jl, jh, level)); }
