// Stand-in of advection.cpp: every function of advection.h.
#include "stand_in.h"

void advect_uv_velocity(double, int, int, int, int) { STAND_IN("advect_uv_velocity"); }    // advection.cpp:269
void advect_uvw_velocity(double, int, int, int, int) { STAND_IN("advect_uvw_velocity"); }  // :304
void advect_theta(double, int, int, int, int) { STAND_IN("advect_theta"); }                // :349
void advect_scalar(double, int, int, int, int, double *, double *, double *, struct cell *) { STAND_IN("advect_scalar"); }  // :383
void advect_scalar(double, int, int, int, int, double *, double *, double *, double *, struct cell *, struct cell *) {
    STAND_IN("advect_scalar");  // :417
}
void advect_microphysics_cell(double, int il, int ih, int jl, int jh) {  // :451
    calculate_eulerian_fall_speed_rain(nullptr, nullptr, il, ih, jl, jh);
    STAND_IN("advect_microphysics_cell");
}
void advect_ice_cell(double, int, int, int, int) { STAND_IN("advect_ice_cell"); }  // :508
void advect_qv(double, int, int, int, int) { STAND_IN("advect_qv"); }              // :577
