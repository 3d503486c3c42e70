// Stand-in of pressure.cpp: the two entry points of pressure.h:7-8.
#include "stand_in.h"

void initialize_pressure_solver() { STAND_IN("initialize_pressure_solver"); }  // pressure.cpp:99
void solve_pressure(double) { STAND_IN("solve_pressure"); }                     // pressure.cpp:142
