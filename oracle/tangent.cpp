// TEST INFRASTRUCTURE ONLY: the tangent-linear model of the oracle.
//
// The oracle's own source (dyncore_oracle.c) is compiled again with every FP64 quantity carrying a
// tangent (forward-mode dual numbers), so each operation is differentiated exactly as the oracle
// evaluates it: the result is the derivative of the oracle's discrete map, with every branch (upwind
// signs, Kessler thresholds, clamps) taken as the base state takes it.  That is the map the HIP
// adjoint transposes ("upwind signs held at the recorded trajectory's values"), so for any direction
// d the adjoint gradient g must satisfy <g, d> = (tangent of J along d) to rounding.  With every basis
// direction the tangent gives the reference gradient itself, independently of the HIP adjoint code.
//
// Exported (C ABI, liboracle_tlm.so): orc_tlm_create / destroy / step / error, and fields and columns
// as (value, tangent) pairs of double arrays:
//   orc_tlm_set_field(s, field, role, value, tangent)   tangent may be NULL (zero)
//   orc_tlm_get_field(s, field, role, value, tangent)   either output may be NULL
//   orc_tlm_set_column(s, column, value)                columns carry no tangent
// (The oracle's own entry points are compiled in as well, with dual-number arguments; they are not
// meant to be called from outside.)
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <cstdint>

struct orc_dual {
    double v, d;
    orc_dual() = default;
    orc_dual(double x) : v(x), d(0.0) {}  // NOLINT: constants enter with a zero tangent
    orc_dual(double x, double t) : v(x), d(t) {}
    explicit operator double() const { return v; }
    orc_dual &operator+=(const orc_dual &b) { d += b.d; v += b.v; return *this; }
    orc_dual &operator-=(const orc_dual &b) { d -= b.d; v -= b.v; return *this; }
    orc_dual &operator*=(const orc_dual &b) { d = d * b.v + v * b.d; v *= b.v; return *this; }
    orc_dual &operator/=(const orc_dual &b) { d = (d * b.v - v * b.d) / (b.v * b.v); v /= b.v; return *this; }
};
typedef orc_dual D;
static inline D operator+(D a, const D &b) { return a += b; }
static inline D operator-(D a, const D &b) { return a -= b; }
static inline D operator*(D a, const D &b) { return a *= b; }
static inline D operator/(D a, const D &b) { return a /= b; }
static inline D operator-(const D &a) { return D(-a.v, -a.d); }
static inline D operator+(const D &a) { return a; }
#define DUAL_CMP(op)                                                                    \
    static inline bool operator op(const D &a, const D &b) { return a.v op b.v; }       \
    static inline bool operator op(const D &a, double b) { return a.v op b; }           \
    static inline bool operator op(double a, const D &b) { return a op b.v; }
DUAL_CMP(<)
DUAL_CMP(>)
DUAL_CMP(<=)
DUAL_CMP(>=)
DUAL_CMP(==)
DUAL_CMP(!=)
#undef DUAL_CMP
static inline D sqrt(const D &a) { double r = ::sqrt(a.v); return D(r, a.d * 0.5 / r); }
static inline D exp(const D &a) { double e = ::exp(a.v); return D(e, a.d * e); }
static inline D log(const D &a) { return D(::log(a.v), a.d / a.v); }
static inline D cos(const D &a) { return D(::cos(a.v), -a.d * ::sin(a.v)); }
static inline D sin(const D &a) { return D(::sin(a.v), a.d * ::cos(a.v)); }
static inline D fabs(const D &a) { return a.v < 0 ? -a : a; }
static inline D fmax(const D &a, const D &b) { return (a.v >= b.v || b.v != b.v) ? a : b; }
static inline D fmin(const D &a, const D &b) { return (a.v <= b.v || b.v != b.v) ? a : b; }
static inline D pow(const D &a, double p) {
    double r = ::pow(a.v, p);
    return D(r, a.d == 0.0 ? 0.0 : a.d * p * ::pow(a.v, p - 1.0));
}
static inline D pow(const D &a, const D &p) {
    double r = ::pow(a.v, p.v);
    double dr = (a.d == 0.0 ? 0.0 : a.d * p.v * ::pow(a.v, p.v - 1.0)) + (p.d == 0.0 ? 0.0 : p.d * r * ::log(a.v));
    return D(r, dr);
}

// the grid as the oracle's orc_grid lays it out, with plain doubles (the API below takes this one)
struct orc_grid_f64 {
    int nx, ny, nz;
    double dx, dz, dt;
    int microphysics, rayleigh_level, sponge_north, sponge_south;
};

// The oracle with double meaning a dual number.  System headers were included above, so their
// include guards keep them out of this substitution.
#define double orc_dual
#include "dyncore_oracle.c"
#undef double

extern "C" {

int orc_tlm_set_field(orc_state *s, int f, int role, const double *value, const double *tangent) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    orc_dual *tmp = (orc_dual *)malloc(sizeof(orc_dual) * n);
    if (!tmp) return -1;
    for (size_t i = 0; i < n; i++) tmp[i] = orc_dual(value[i], tangent ? tangent[i] : 0.0);
    orc_set_field(s, f, role, tmp);
    free(tmp);
    return 0;
}

int orc_tlm_get_field(orc_state *s, int f, int role, double *value, double *tangent) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    orc_dual *tmp = (orc_dual *)malloc(sizeof(orc_dual) * n);
    if (!tmp) return -1;
    orc_get_field(s, f, role, tmp);
    for (size_t i = 0; i < n; i++) {
        if (value) value[i] = tmp[i].v;
        if (tangent) tangent[i] = tmp[i].d;
    }
    free(tmp);
    return 0;
}

int orc_tlm_set_column(orc_state *s, int c, const double *value) {
    size_t n = c == C_F ? (size_t)s->NY : (size_t)s->NZ;
    orc_dual *tmp = (orc_dual *)malloc(sizeof(orc_dual) * n);
    if (!tmp) return -1;
    for (size_t i = 0; i < n; i++) tmp[i] = orc_dual(value[i]);
    orc_set_column(s, c, tmp);
    free(tmp);
    return 0;
}

orc_state *orc_tlm_create(const orc_grid_f64 *g) {
    orc_grid d;
    d.nx = g->nx;
    d.ny = g->ny;
    d.nz = g->nz;
    d.dx = g->dx;
    d.dz = g->dz;
    d.dt = g->dt;
    d.microphysics = g->microphysics;
    d.rayleigh_level = g->rayleigh_level;
    d.sponge_north = g->sponge_north;
    d.sponge_south = g->sponge_south;
    return orc_create(&d);
}
void orc_tlm_destroy(orc_state *s) { orc_destroy(s); }
void orc_tlm_step(orc_state *s, int n) { orc_step(s, n); }
int orc_tlm_error(const orc_state *s) { return orc_error(s); }
}
