// Stand-in of microphysics.cpp: the kept set-up, precipitation and fall-speed functions, and the
// entry points the drop-in library replaces.
#include "stand_in.h"

void init_microphysics(int, int) { init_calls++; }  // microphysics.cpp:41 (kept)
void precip_rate(int, int, int, int, double *, double *, double *) {}  // microphysics.cpp:143 (kept)
void calculate_eulerian_fall_speed_rain(double *vt, double *qr, int il, int ih, int jl, int jh) {  // :169 (kept)
    calculate_rain_fall_speed_kessler(vt, qr, il, ih, jl, jh);
}
void run_microphysics(int, int, int, int) { STAND_IN("run_microphysics"); }                     // :97
void microphysics_advance_inner(size_t) { STAND_IN("microphysics_advance_inner"); }             // :642
void microphysics_advance(size_t) { STAND_IN("microphysics_advance"); }                         // :657
void zero_moisture(int, int, int, int, int) { STAND_IN("zero_moisture"); }                      // :675
