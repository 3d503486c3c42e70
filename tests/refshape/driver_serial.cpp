// Stand-in of driver_serial.cpp: the reference's serial initialisation order (initialize_serial,
// driver_serial.cpp:44-77), its RK3 step (s_integrate_rk3, driver_serial.cpp:178-286) calling the
// reference entry points by name, and a main that runs it on a state read from a file.
//
//   refshape_driver <input> <output> <steps>   run and write every field (all roles)
//   refshape_driver --load-only                load the libraries and exit (symbol-binding checks)
#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "../../include/dyncore_refapi.h"
#include "stand_in.h"

extern int NX, NY, NZ, USE_MICROPHYSICS, USE_TURBULENT_STRESS, USE_EXPLICIT_DIFFUSION, raydampheight;
extern int iebuffer, iwbuffer, jnbuffer, jsbuffer;
int load_input(const char *path);
int write_output(const char *path);

extern double *zu;

void initialize_serial() {  // driver_serial.cpp:44-77, with the initial state already in the globals
    if (USE_TURBULENT_STRESS) init_kmix(NX, NY, NZ, zu);
    if (USE_MICROPHYSICS) init_microphysics(NX, NY);
    initialize_landsea(nullptr);
    initialize_pressure_solver();
    init_boundaries(iebuffer, iwbuffer, jnbuffer, jsbuffer, -1);
    init_damping(NX, NY, NZ);
}

void s_integrate_rk3() {  // driver_serial.cpp:178-286 (no explicit diffusion, no terrain)
    double steps[] = {1. / 3., 0.5, 1.0};
    int size = NX * NY * NZ;
    if (USE_TURBULENT_STRESS) calculate_diff_tend(1, NX - 1, 1, NY - 1);
    for (int s = 0; s < 3; s++) {
        compute_sign_cells(0, NX, 0, NY);
        integrate_non_hydro(steps[s], 0, NX, 0, NY);
        advect_theta(steps[s], 3, NX - 3, 3, NY - 3);
        if (USE_TURBULENT_STRESS) apply_scalar_diffusion(steps[s], 1, NX - 1, 1, NY - 1);
        if (USE_MICROPHYSICS) {
            advect_microphysics_cell(steps[s], 3, NX - 3, 3, NY - 3);
            if (USE_TURBULENT_STRESS) apply_moisture_diffusion(steps[s], 1, NX - 1, 1, NY - 1);
            zero_moisture(1, NX - 1, 1, NY - 1, size);
        }
        apply_boundary_condition(0);
        if (USE_MICROPHYSICS) {
            apply_boundary_condition_microphysics(0);
            if (s < 2) microphysics_advance_inner(size * sizeof(double));
        }
        if (s < 2) advance_inner(size * sizeof(double));
    }
    if (USE_MICROPHYSICS) {
        run_microphysics(1, NX - 1, 1, NY - 1);
        apply_boundary_condition_microphysics(0);
        microphysics_advance(size * sizeof(double));
    }
    apply_boundary_condition(0);
    rayleigh_damping(1, NX - 1, 1, NY - 1, raydampheight);
    advance_outer(size * sizeof(double));
}

int main(int argc, char **argv) {
    if (argc == 2 && strcmp(argv[1], "--load-only") == 0) return 0;
    if (argc != 4) {
        fprintf(stderr, "usage: %s <input> <output> <steps> | --load-only\n", argv[0]);
        return 2;
    }
    if (load_input(argv[1])) return 2;
    dyn_refapi_exit_on_error(0);
    initialize_serial();
    int steps = atoi(argv[3]);
    for (int n = 0; n < steps && !dyn_refapi_error()[0]; n++) {  // s_run_model, driver_serial.cpp:310-372
        s_integrate_rk3();
        bigcounter++;
    }
    if (!dyn_refapi_error()[0]) dyn_refapi_download();
    if (dyn_refapi_error()[0]) {
        fprintf(stderr, "drop-in error: %s\n", dyn_refapi_error());
        return 4;
    }
    if (init_calls != 3 + (USE_MICROPHYSICS ? 1 : 0) + (USE_TURBULENT_STRESS ? 1 : 0)) {
        fprintf(stderr, "kept initialisers ran %d times\n", init_calls);
        return 5;
    }
    dyn_refapi_reset();
    return write_output(argv[2]) ? 2 : 0;
}
