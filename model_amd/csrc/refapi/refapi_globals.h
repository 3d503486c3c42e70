// Declarations of the reference's host globals that its dynamical-core entry points use
// (model.h:16-26, 63-91, 220-245; microphysics.h:18-25; stdafx.h:51-68; boundaries.h:4-7; damping.h:2;
// turbulence.h:2).
// In the reference build these are defined by initializer.cpp; the standalone build defines them in
// refapi_host.cpp.
#pragma once

extern int NX, NY, NZ, NYNZ;
extern double dx, dy, dz, dt;
extern double dtx, dty, dtz, one_d_dx, one_d_dy, one_d_dz;
extern double *us, *vs, *ws, *ths;
extern double *ups, *vps, *wps, *thps;
extern double *ums, *vms, *wms, *thms;
extern double *pis;
extern double *m_ubar, *m_vbar, *m_wbar, *m_thbar, *m_qbar, *m_pbar;
extern double *f;
extern double *zu, *zw, *rhou, *rhow, *tb, *tbv, *tbw, *qb, *pib, *one_d_rhou, *one_d_rhow, *zsu, *zsw, *mu, *mw;
extern double *qvps, *qvs, *qvms, *qcps, *qcs, *qcms, *qrps, *qrs, *qrms;
extern double *vts, *accRain;
extern int MICROPHYSICS_OPTION, USE_MICROPHYSICS, USE_ICE, RAIN_FALLOUT, PERIODIC_BOUNDARIES;
extern int USE_TURBULENT_STRESS, USE_EXPLICIT_DIFFUSION;
extern int iwbuffer, iebuffer, jnbuffer, jsbuffer;
extern int raydampheight;
extern double *landsea;  // turbulence.cpp:32, filled by initialize_landsea (NX*NY, < 0.5 land)
