// Stand-in of damping.cpp: init_damping (kept; driver_serial.cpp:72 calls it) and rayleigh_damping.
#include "stand_in.h"

void init_damping(int, int, int) { init_calls++; }                               // damping.cpp:101 (kept)
void rayleigh_damping(int, int, int, int, int) { STAND_IN("rayleigh_damping"); }  // damping.cpp:1148
