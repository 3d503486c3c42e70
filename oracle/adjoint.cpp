// TEST INFRASTRUCTURE ONLY: the reverse-mode (discrete adjoint) build of the oracle.
//
// The oracle's own source (dyncore_oracle.c) is compiled again with every FP64 quantity a tape-backed
// value: each arithmetic operation on values that depend on the inputs appends one node to a tape with
// its operands and its two partial derivatives, evaluated at the operation's own arguments.  Sweeping
// the tape backwards from a seed on the final fields gives the gradient of the seeded scalar with
// respect to the initial fields: the discrete adjoint of the oracle's map, in FP64, with every branch
// (upwind signs, thresholds, clamps) taken as the base trajectory takes it.  This is the CPU reference
// the HIP adjoint (dyn_adjoint_backward) is held to; it shares no code with it.
//
// Exported (C ABI, liboracle_adj.so), mirroring dyn_adjoint_backward:
//   orc_adj_create / destroy / step / error / set_column / set_field / get_field
//   orc_adj_set_initial(s, field, value)   value set into M, current and forecast of a prognostic, each
//                                          element one independent input node
//   orc_adj_gradient(s, seed_u, seed_v, seed_w, seed_th, grad_u, grad_v, grad_w, grad_th)
//       dJ/d(initial field) for J = sum(seed * final current field); NULL seeds are zero, NULL
//       gradients are skipped
//   orc_adj_tape_size()                    nodes recorded so far
// One tape per process: create, set the initial fields, step and take gradients before the next state.
#include <math.h>
#include <stddef.h>
#include <stdlib.h>
#include <string.h>

#include <cstdint>
#include <vector>

namespace {
// node 0 stands for every value that does not depend on the inputs
struct Tape {
    std::vector<int32_t> a, b;
    std::vector<double> pa, pb;
    Tape() { clear(); }
    void clear() {
        a.assign(1, 0);
        b.assign(1, 0);
        pa.assign(1, 0.0);
        pb.assign(1, 0.0);
    }
    int32_t push(int32_t ia, double xa, int32_t ib, double xb) {
        a.push_back(ia);
        b.push_back(ib);
        pa.push_back(xa);
        pb.push_back(xb);
        return (int32_t)(a.size() - 1);
    }
};
Tape g_tape;
}  // namespace

struct orc_adouble {
    double v;
    int32_t id;
    orc_adouble() = default;
    orc_adouble(double x) : v(x), id(0) {}  // NOLINT: constants enter as passive values
    explicit operator double() const { return v; }
};
typedef orc_adouble A;

static inline A mk(double v, int32_t ia, double xa, int32_t ib, double xb) {
    A r;
    r.v = v;
    r.id = (ia || ib) ? g_tape.push(ia, xa, ib, xb) : 0;
    return r;
}
static inline A operator+(const A &x, const A &y) { return mk(x.v + y.v, x.id, 1.0, y.id, 1.0); }
static inline A operator-(const A &x, const A &y) { return mk(x.v - y.v, x.id, 1.0, y.id, -1.0); }
static inline A operator*(const A &x, const A &y) { return mk(x.v * y.v, x.id, y.v, y.id, x.v); }
static inline A operator/(const A &x, const A &y) {
    const double r = x.v / y.v;
    return mk(r, x.id, 1.0 / y.v, y.id, -r / y.v);
}
static inline A operator-(const A &x) { return mk(-x.v, x.id, -1.0, 0, 0.0); }
static inline A operator+(const A &x) { return x; }
static inline A &operator+=(A &x, const A &y) { return x = x + y; }
static inline A &operator-=(A &x, const A &y) { return x = x - y; }
static inline A &operator*=(A &x, const A &y) { return x = x * y; }
static inline A &operator/=(A &x, const A &y) { return x = x / y; }
#define ADJ_CMP(op)                                                                   \
    static inline bool operator op(const A &x, const A &y) { return x.v op y.v; }     \
    static inline bool operator op(const A &x, double y) { return x.v op y; }         \
    static inline bool operator op(double x, const A &y) { return x op y.v; }
ADJ_CMP(<)
ADJ_CMP(>)
ADJ_CMP(<=)
ADJ_CMP(>=)
ADJ_CMP(==)
ADJ_CMP(!=)
#undef ADJ_CMP
static inline A sqrt(const A &x) {
    const double r = ::sqrt(x.v);
    return mk(r, x.id, 0.5 / r, 0, 0.0);
}
static inline A exp(const A &x) {
    const double e = ::exp(x.v);
    return mk(e, x.id, e, 0, 0.0);
}
static inline A log(const A &x) { return mk(::log(x.v), x.id, 1.0 / x.v, 0, 0.0); }
static inline A cos(const A &x) { return mk(::cos(x.v), x.id, -::sin(x.v), 0, 0.0); }
static inline A sin(const A &x) { return mk(::sin(x.v), x.id, ::cos(x.v), 0, 0.0); }
static inline A fabs(const A &x) { return x.v < 0 ? -x : x; }
static inline A fmax(const A &x, const A &y) { return (x.v >= y.v || y.v != y.v) ? x : y; }
static inline A fmin(const A &x, const A &y) { return (x.v <= y.v || y.v != y.v) ? x : y; }
static inline A pow(const A &x, double p) {
    const double r = ::pow(x.v, p);
    return mk(r, x.id, x.id ? p * ::pow(x.v, p - 1.0) : 0.0, 0, 0.0);
}
static inline A pow(const A &x, const A &p) {
    const double r = ::pow(x.v, p.v);
    return mk(r, x.id, x.id ? p.v * ::pow(x.v, p.v - 1.0) : 0.0, p.id, p.id ? r * ::log(x.v) : 0.0);
}

// the grid as the oracle's orc_grid lays it out, with plain doubles (the API below takes this one)
struct orc_grid_f64 {
    int nx, ny, nz;
    double dx, dz, dt;
    int microphysics, rayleigh_level, sponge_north, sponge_south;
};

// The oracle with double meaning a tape-backed value.  System headers were included above, so their
// include guards keep them out of this substitution.
#define double orc_adouble
#include "dyncore_oracle.c"
#undef double

namespace {
// input node ids of the initial u, v, w, theta (one per element), filled by orc_adj_set_initial
std::vector<int32_t> g_inputs[4];
}  // namespace

extern "C" {

orc_state *orc_adj_create(const orc_grid_f64 *g) {
    g_tape.clear();
    for (auto &v : g_inputs) v.clear();
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

void orc_adj_destroy(orc_state *s) {
    orc_destroy(s);
    g_tape.clear();
    for (auto &v : g_inputs) v.clear();
}

int orc_adj_set_column(orc_state *s, int c, const double *value) {
    size_t n = c == C_F ? (size_t)s->NY : (size_t)s->NZ;
    std::vector<orc_adouble> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = orc_adouble(value[i]);
    orc_set_column(s, c, tmp.data());
    return 0;
}

int orc_adj_set_field(orc_state *s, int f, int role, const double *value) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    std::vector<orc_adouble> tmp(n);
    for (size_t i = 0; i < n; i++) tmp[i] = orc_adouble(value[i]);
    orc_set_field(s, f, role, tmp.data());
    return 0;
}

int orc_adj_set_initial(orc_state *s, int f, const double *value) {
    if (f < F_U || f > F_TH) return -1;
    size_t n = (size_t)s->NX * s->NYNZ;
    std::vector<orc_adouble> tmp(n);
    std::vector<int32_t> &ids = g_inputs[f];
    ids.resize(n);
    for (size_t i = 0; i < n; i++) {
        tmp[i].v = value[i];
        tmp[i].id = ids[i] = g_tape.push(0, 0.0, 0, 0.0);
    }
    for (int role = R_M; role <= R_P; role++) orc_set_field(s, f, role, tmp.data());
    return 0;
}

int orc_adj_get_field(orc_state *s, int f, int role, double *value) {
    size_t n = f == F_ACCRAIN ? (size_t)s->NX * s->NY : (size_t)s->NX * s->NYNZ;
    std::vector<orc_adouble> tmp(n);
    orc_get_field(s, f, role, tmp.data());
    for (size_t i = 0; i < n; i++) value[i] = tmp[i].v;
    return 0;
}

void orc_adj_step(orc_state *s, int n) { orc_step(s, n); }
int orc_adj_error(const orc_state *s) { return orc_error(s); }
long long orc_adj_tape_size(void) { return (long long)g_tape.a.size(); }

int orc_adj_gradient(orc_state *s, const double *seed_u, const double *seed_v, const double *seed_w, const double *seed_th, double *grad_u,
                     double *grad_v, double *grad_w, double *grad_th) {
    const size_t n = (size_t)s->NX * s->NYNZ, nodes = g_tape.a.size();
    std::vector<double> adj(nodes, 0.0);
    const double *seeds[4] = {seed_u, seed_v, seed_w, seed_th};
    std::vector<orc_adouble> fin(n);
    for (int f = F_U; f <= F_TH; f++) {
        if (!seeds[f]) continue;
        orc_get_field(s, f, R_CUR, fin.data());
        for (size_t i = 0; i < n; i++)
            if (fin[i].id) adj[fin[i].id] += seeds[f][i];
    }
    for (size_t k = nodes - 1; k > 0; k--) {
        const double w = adj[k];
        if (w == 0.0) continue;
        if (g_tape.a[k]) adj[g_tape.a[k]] += g_tape.pa[k] * w;
        if (g_tape.b[k]) adj[g_tape.b[k]] += g_tape.pb[k] * w;
    }
    double *grads[4] = {grad_u, grad_v, grad_w, grad_th};
    for (int f = F_U; f <= F_TH; f++) {
        if (!grads[f]) continue;
        if (g_inputs[f].size() != n) return -1;
        for (size_t i = 0; i < n; i++) grads[f][i] = adj[g_inputs[f][i]];
    }
    return 0;
}
}
