// Stand-in of kessler.cpp: the fall speed (kept; microphysics.cpp:173 calls it) and the scheme.
#include "stand_in.h"

void calculate_rain_fall_speed_kessler(double *, double *, int, int, int, int) {}  // kessler.cpp:222 (kept)
void run_kessler_microphysics(int il, int ih, int jl, int jh) {                    // kessler.cpp:25
    precip_rate(il, ih, jl, jh, nullptr, nullptr, nullptr);
    STAND_IN("run_kessler_microphysics");
}
