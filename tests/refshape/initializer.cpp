// Stand-in of initializer.cpp: it defines the model's host globals (initializer.cpp:24-144) with the
// reference's types, and reads them from a file the test writes (the reference fills them from its
// input parameters and initial-condition code).
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/dyncore.h"
#include "stand_in.h"

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
int raydampheight = 0;
int bigcounter = 0;
int init_calls = 0;
double *landsea_input = nullptr;
extern int jnbuffer, jsbuffer;

double *field_ptr(int field, int role) {
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

static double **column_slot(int c) {
    static double **cols[DYN_COLUMN_COUNT] = {&zu, &zw, &zsu, &zsw, &mu, &mw, &rhou, &rhow, &one_d_rhou, &one_d_rhow,
                                              &tb, &tbv, &tbw, &qb, &pib, &f};
    return cols[c];
}

static void rd(void *p, size_t bytes, FILE *fp) {
    if (fread(p, 1, bytes, fp) != bytes) {
        fprintf(stderr, "short input file\n");
        exit(2);
    }
}

// File layout: int32 {nx, ny, nz, microphysics, raydampheight, jn, js, turbulence, mask}, float64 {dx,
// dz, dt}, the 16 columns (f has NY values, the rest NZ), the land-sea mask (NX*NY doubles) when mask
// is 1, then for every field in dyn_field order its roles (three for a prognostic, one otherwise) as
// NX*NY*NZ doubles (accrain NX*NY).
int load_input(const char *path) {
    FILE *fp = fopen(path, "rb");
    if (!fp) return 1;
    int hdr[9];
    double d[3];
    rd(hdr, sizeof hdr, fp);
    rd(d, sizeof d, fp);
    NX = hdr[0];
    NY = hdr[1];
    NZ = hdr[2];
    NYNZ = NY * NZ;
    MICROPHYSICS_OPTION = hdr[3];
    USE_MICROPHYSICS = hdr[3] > 0;
    raydampheight = hdr[4];
    jnbuffer = hdr[5];
    jsbuffer = hdr[6];
    USE_TURBULENT_STRESS = hdr[7];
    dx = dy = d[0];
    dz = d[1];
    dt = d[2];
    dtx = dt / dx;
    dty = dt / dy;
    dtz = dt / dz;
    one_d_dx = 1.0 / dx;
    one_d_dy = 1.0 / dy;
    one_d_dz = 1.0 / dz;
    // allocate zero-filled, as initialize_subarray does (initializer.cpp:421-485)
    size_t n = (size_t)NX * NY * NZ;
    double **fields[] = {&us, &vs, &ws, &ths, &ups, &vps, &wps, &thps, &ums, &vms, &wms, &thms, &pis, &m_ubar, &m_vbar, &m_wbar,
                         &m_thbar, &m_qbar, &m_pbar, &qvps, &qvs, &qvms, &qcps, &qcs, &qcms, &qrps, &qrs, &qrms, &vts};
    for (double **p : fields) *p = (double *)calloc(n, sizeof(double));
    accRain = (double *)calloc((size_t)NX * NY, sizeof(double));
    for (int c = 0; c < DYN_COLUMN_COUNT; c++) {
        size_t len = c == DYN_COL_F ? (size_t)NY : (size_t)NZ;
        *column_slot(c) = (double *)calloc(len, sizeof(double));
        rd(*column_slot(c), len * sizeof(double), fp);
    }
    if (hdr[8]) {
        landsea_input = (double *)malloc((size_t)NX * NY * sizeof(double));
        rd(landsea_input, (size_t)NX * NY * sizeof(double), fp);
    }
    for (int fld = 0; fld < DYN_FIELD_COUNT; fld++) {
        int roles = fld <= DYN_QR ? 3 : 1;
        size_t len = fld == DYN_ACCRAIN ? (size_t)NX * NY : n;
        for (int r = 0; r < roles; r++) rd(field_ptr(fld, r), len * sizeof(double), fp);
    }
    fclose(fp);
    return 0;
}

int write_output(const char *path) {
    FILE *fp = fopen(path, "wb");
    if (!fp) return 1;
    size_t n = (size_t)NX * NY * NZ;
    for (int fld = 0; fld < DYN_FIELD_COUNT; fld++) {
        int roles = fld <= DYN_QR ? 3 : 1;
        size_t len = fld == DYN_ACCRAIN ? (size_t)NX * NY : n;
        for (int r = 0; r < roles; r++)
            if (fwrite(field_ptr(fld, r), sizeof(double), len, fp) != len) return 1;
    }
    fclose(fp);
    return 0;
}
