// Test executable shaped like the reference's serial build: one source file per reference object,
// each defining what that object defines.  Functions the drop-in library replaces have CPU stand-in
// bodies that stop the program, so a run that reaches one shows a binding to the wrong definition.
// Functions the library does not replace (init_boundaries, precip_rate, ...) are ordinary code.
#pragma once
#include <cstddef>
#include <cstdio>
#include <cstdlib>

#define STAND_IN(name)                                                                                 \
    do {                                                                                               \
        fprintf(stderr, "CPU stand-in %s ran: the call did not bind to libdyncore_refapi\n", name);    \
        exit(3);                                                                                       \
    } while (0)

struct cell;

// kept-object functions other kept objects call (declared as the reference's headers declare them)
void precip_rate(int il, int ih, int jl, int jh, double *vel, double *hydro_field, double *output);  // microphysics.h:80
void calculate_rain_fall_speed_kessler(double *vt, double *qr, int il, int ih, int jl, int jh);    // microphysics.h:81
void calculate_eulerian_fall_speed_rain(double *vt, double *qr, int il, int ih, int jl, int jh);   // microphysics.h:85
void init_boundaries(int eastLength, int westLength, int northLength, int southLength, int halo);   // boundaries.h:16
void init_damping(int ni, int nj, int nk);                                                         // damping.h:14
void init_microphysics(int nx, int ny);                                                            // microphysics.h:66
void switch_array(double **var1, double **var2);                                                   // solver.h:7
void init_kmix(int nx, int ny, int nz, double *zlevs);                                             // turbulence.h:7
void initialize_landsea(const char *name);                                                         // turbulence.h:6
extern int bigcounter;
extern int init_calls;  // counts the kept initialisers that ran
