// s_integrate_rk3 of driver_serial.cpp:178-286, restated over the drop-in entry points, and the
// standalone host globals (what initializer.cpp defines in the reference build) with the hooks the
// tests use to fill them.
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../../include/dyncore.h"
#include "../../../include/dyncore_refapi.h"
#include "refapi_globals.h"

// ---- host globals (initializer.cpp:24-144, boundaries.cpp:23-26, microphysics switches) -------------
int NX, NY, NZ, NYNZ;
double dx, dy, dz, dt;
double dtx, dty, dtz, one_d_dx, one_d_dy, one_d_dz;
double *us, *vs, *ws, *ths;
double *ups, *vps, *wps, *thps;
double *ums, *vms, *wms, *thms;
double *pis;
double *m_ubar, *m_vbar, *m_wbar, *m_thbar, *m_qbar, *m_pbar;
double *f;
double *zu, *zw, *rhou, *rhow, *tb, *tbv, *tbw, *qb, *pib, *one_d_rhou, *one_d_rhow, *zsu, *zsw, *mu, *mw;
double *qvps, *qvs, *qvms, *qcps, *qcs, *qcms, *qrps, *qrs, *qrms;
double *vts, *accRain;
int MICROPHYSICS_OPTION = 0, USE_MICROPHYSICS = 0, USE_ICE = 0, RAIN_FALLOUT = 1, PERIODIC_BOUNDARIES = 1;
int USE_TURBULENT_STRESS = 0, USE_EXPLICIT_DIFFUSION = 0;
int iwbuffer = 5, iebuffer = 5, jnbuffer = 5, jsbuffer = 5;
int raydampheight = 0;

// ---- the serial RK3 step, driver_serial.cpp:178-286 --------------------------------------------------
void s_integrate_rk3() {
    double steps[] = {1. / 3., 0.5, 1.0};  // driver_serial.cpp:180
    int size = NX * NY * NZ;
    if (USE_EXPLICIT_DIFFUSION || USE_TURBULENT_STRESS) {
        printf("explicit diffusion and turbulence are outside the implemented path\n");
        exit(0);
    }
    for (int s = 0; s < 3; s++) {
        compute_sign_cells(0, NX, 0, NY);
        integrate_non_hydro(steps[s], 0, NX, 0, NY);
        advect_theta(steps[s], 3, NX - 3, 3, NY - 3);
        if (USE_MICROPHYSICS) {
            advect_microphysics_cell(steps[s], 3, NX - 3, 3, NY - 3);
            zero_moisture(1, NX - 1, 1, NY - 1, size);
        }
        apply_boundary_condition(0);
        if (USE_MICROPHYSICS) {
            apply_boundary_condition_microphysics(0);
            if (s < 2) microphysics_advance_inner(NX * NY * NZ * sizeof(double));
        }
        if (s < 2) advance_inner(NX * NY * NZ * sizeof(double));
    }
    if (USE_MICROPHYSICS) {
        run_microphysics(1, NX - 1, 1, NY - 1);
        apply_boundary_condition_microphysics(0);
        microphysics_advance(NX * NY * NZ * sizeof(double));
    }
    apply_boundary_condition(0);
    rayleigh_damping(1, NX - 1, 1, NY - 1, raydampheight);
    advance_outer(NX * NY * NZ * sizeof(double));
}

// ---- standalone set-up hooks (the reference's initializer fills these globals itself) ---------------
extern "C" {

// allocate every global array zero-filled, as initialize_subarray does (initializer.cpp:421-485)
int dyn_refapi_setup(int nx, int ny, int nz, double ddx, double ddz, double ddt, int microphysics, int raylevel, int jn, int js) {
    NX = nx;
    NY = ny;
    NZ = nz;
    NYNZ = ny * nz;
    dx = ddx;
    dy = ddx;
    dz = ddz;
    dt = ddt;
    dtx = dt / dx;
    dty = dt / dy;
    dtz = dt / dz;
    one_d_dx = 1.0 / dx;
    one_d_dy = 1.0 / dy;
    one_d_dz = 1.0 / dz;
    MICROPHYSICS_OPTION = microphysics;
    USE_MICROPHYSICS = microphysics > 0;
    USE_ICE = microphysics == 2;
    raydampheight = raylevel;
    jnbuffer = jn;
    jsbuffer = js;
    size_t n = (size_t)nx * ny * nz;
    double **fields[] = {&us, &vs, &ws, &ths, &ups, &vps, &wps, &thps, &ums, &vms, &wms, &thms, &pis, &m_ubar, &m_vbar, &m_wbar,
                         &m_thbar, &m_qbar, &m_pbar, &qvps, &qvs, &qvms, &qcps, &qcs, &qcms, &qrps, &qrs, &qrms, &vts};
    for (double **p : fields) {
        free(*p);
        *p = (double *)calloc(n, sizeof(double));
    }
    free(accRain);
    accRain = (double *)calloc((size_t)nx * ny, sizeof(double));
    double **cols[] = {&zu, &zw, &zsu, &zsw, &mu, &mw, &rhou, &rhow, &one_d_rhou, &one_d_rhow, &tb, &tbv, &tbw, &qb, &pib};
    for (double **p : cols) {
        free(*p);
        *p = (double *)calloc(nz, sizeof(double));
    }
    free(f);
    f = (double *)calloc(ny, sizeof(double));
    return 0;
}

// host pointer of a global field (role: 0 = M, 1 = current, 2 = forecast) or column
double *dyn_refapi_field(int field, int role) {
    switch (field) {
    case DYN_U: return role == 0 ? ums : role == 1 ? us : ups;
    case DYN_V: return role == 0 ? vms : role == 1 ? vs : vps;
    case DYN_W: return role == 0 ? wms : role == 1 ? ws : wps;
    case DYN_TH: return role == 0 ? thms : role == 1 ? ths : thps;
    case DYN_QV: return role == 0 ? qvms : role == 1 ? qvs : qvps;
    case DYN_QC: return role == 0 ? qcms : role == 1 ? qcs : qcps;
    case DYN_QR: return role == 0 ? qrms : role == 1 ? qrs : qrps;
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

double *dyn_refapi_column(int column) {
    double *cols[DYN_COLUMN_COUNT] = {zu, zw, zsu, zsw, mu, mw, rhou, rhow, one_d_rhou, one_d_rhow, tb, tbv, tbw, qb, pib, f};
    return (column >= 0 && column < DYN_COLUMN_COUNT) ? cols[column] : nullptr;
}

// initialize_serial's solver part (driver_serial.cpp:58-60): solver set-up, then the sponge set-up
int dyn_refapi_initialize(void) {
    initialize_pressure_solver();
    init_boundaries(iebuffer, iwbuffer, jnbuffer, jsbuffer, -1);
    return dyn_refapi_error()[0] ? DYN_E_STATE : DYN_OK;
}

// s_run_model's stepping (driver_serial.cpp:310-372) without output
int dyn_refapi_run(int count) {
    for (int n = 0; n < count; n++) {
        s_integrate_rk3();
        if (dyn_refapi_error()[0]) return DYN_E_STATE;
    }
    return DYN_OK;
}
}
