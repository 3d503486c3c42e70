// Stand-in of turbulence.cpp: the mask and tendency globals (turbulence.cpp:32-35), the kept set-up
// (init_kmix, initialize_landsea) and the four entry points the drop-in library replaces.
#include <cstring>

#include "stand_in.h"

double *landsea;
double *u_friction, *v_friction, *w_friction, *t_diffusion;
double *qv_diffusion, *qc_diffusion, *qr_diffusion;
extern int NX, NY, NZ;
extern double *landsea_input;  // the mask the input file carried (nullptr: none)

void init_kmix(int nx, int ny, int nz, double *) {  // turbulence.cpp:85 (kept): host tendency arrays
    size_t n = (size_t)nx * ny * nz;
    double **arrays[] = {&u_friction, &v_friction, &w_friction, &t_diffusion, &qv_diffusion, &qc_diffusion, &qr_diffusion};
    for (double **a : arrays) *a = (double *)calloc(n, sizeof(double));
    init_calls++;
}
// turbulence.cpp:297 (kept): the mask from a file when one is named, all ocean (1.0) otherwise; here
// the "file" is the mask the input carried
void initialize_landsea(const char *) {
    size_t n = (size_t)NX * NY;
    landsea = (double *)malloc(n * sizeof(double));
    for (size_t p = 0; p < n; p++) landsea[p] = landsea_input ? landsea_input[p] : 1.0;
    init_calls++;
}
void calculate_diff_tend(int, int, int, int) { STAND_IN("calculate_diff_tend"); }                       // :158
void apply_velocity_diffusion(double, int, int, int, int) { STAND_IN("apply_velocity_diffusion"); }     // :233
void apply_scalar_diffusion(double, int, int, int, int) { STAND_IN("apply_scalar_diffusion"); }         // :263
void apply_moisture_diffusion(double, int, int, int, int) { STAND_IN("apply_moisture_diffusion"); }     // :278
