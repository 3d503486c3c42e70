// Stand-in of boundaries.cpp: the sponge widths (boundaries.cpp:23-26), init_boundaries (kept) and the
// two boundary entry points the drop-in library replaces.
#include "stand_in.h"

int iwbuffer = 5;
int iebuffer = 5;
int jnbuffer = 5;
int jsbuffer = 5;

void init_boundaries(int, int, int, int, int) { init_calls++; }  // boundaries.cpp:44 (kept)
void apply_boundary_condition(int) { STAND_IN("apply_boundary_condition"); }                    // :119
void apply_boundary_condition_microphysics(int) { STAND_IN("apply_boundary_condition_microphysics"); }  // :200
