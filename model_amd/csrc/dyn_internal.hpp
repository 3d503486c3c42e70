// Internal definitions shared by the HIP translation units of libdyncore.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/dyncore.h"

namespace dyn {

// physical constants, model.h:250-258
constexpr double kRd = 287.0;
constexpr double kCp = 1004;
constexpr double kP0 = 100000;
constexpr double kGrav = 9.8;
constexpr double kLv = 2.5e6;
constexpr double kTrigPi = 3.14159265358979324;

void set_error(const std::string &msg);

// Face interpolants of fluxes.cpp:110-206 on explicit operands (p0 = the value on the low side of the
// face, sign = +-1 from the advecting velocity): the 5th-order upwind-biased one, c6 + sign d5 / 60, and
// the 3rd-order one.  Every stencil kernel takes them from here, so the fused and per-operator forms of
// an operator agree to the last bit.  FP64 evaluates the reference's expressions in their order; FP32
// evaluates the same polynomials with fused multiply-adds (13 and 9 operations instead of 18 and 12,
// one rounding less per product): the stage kernels are bound by instruction issue.
template <typename T>
__device__ __forceinline__ T face_i5(T pm2, T pm1, T p0, T p1, T p2, T p3, T sign) {
    const T c37_60 = T(37.0 / 60.0), c2_15 = T(2.0 / 15.0), c1_60 = T(1.0 / 60.0);
    T c6 = c37_60 * (p1 + p0) - c2_15 * (p2 + pm1) + c1_60 * (p3 + pm2);
    T d5 = -(p3 - pm2) + T(5.0) * (p2 - pm1) - T(10.0) * (p1 - p0);
    return c6 + c1_60 * sign * d5;
}
template <>
__device__ __forceinline__ float face_i5<float>(float pm2, float pm1, float p0, float p1, float p2, float p3, float sign) {
    const float c37_60 = float(37.0 / 60.0), c2_15 = float(2.0 / 15.0), c1_60 = float(1.0 / 60.0);
    const float c6 = __builtin_fmaf(c37_60, p1 + p0, __builtin_fmaf(-c2_15, p2 + pm1, c1_60 * (p3 + pm2)));
    const float d5 = __builtin_fmaf(5.0f, p2 - pm1, __builtin_fmaf(-10.0f, p1 - p0, pm2 - p3));
    return __builtin_fmaf(c1_60 * sign, d5, c6);
}
template <typename T>
__device__ __forceinline__ T face_i3(T pm1, T p0, T p1, T p2, T sign) {
    const T c7_12 = T(7.0 / 12.0), c1_12 = T(1.0 / 12.0);
    return (c7_12 * (p1 + p0) - c1_12 * (p2 + pm1)) + c1_12 * sign * ((p2 - pm1) - T(3.0) * (p1 - p0));
}
template <>
__device__ __forceinline__ float face_i3<float>(float pm1, float p0, float p1, float p2, float sign) {
    const float c7_12 = float(7.0 / 12.0), c1_12 = float(1.0 / 12.0);
    const float c4 = __builtin_fmaf(c7_12, p1 + p0, -(c1_12 * (p2 + pm1)));
    return __builtin_fmaf(c1_12 * sign, __builtin_fmaf(-3.0f, p1 - p0, p2 - pm1), c4);
}

#define DYN_HIP(call)                                                                              \
    do {                                                                                           \
        hipError_t _e = (call);                                                                    \
        if (_e != hipSuccess) {                                                                    \
            ::dyn::set_error(std::string(#call " failed: ") + hipGetErrorString(_e));            \
            return DYN_E_HIP;                                                                      \
        }                                                                                          \
    } while (0)

#define DYN_LAUNCH_CHECK()                                                                         \
    do {                                                                                           \
        hipError_t _e = hipGetLastError();                                                         \
        if (_e != hipSuccess) {                                                                    \
            ::dyn::set_error(std::string("kernel launch failed: ") + hipGetErrorString(_e));     \
            return DYN_E_HIP;                                                                      \
        }                                                                                          \
    } while (0)

// Grid geometry and the per-level constants every kernel needs, passed by value.
struct Geo {
    int NX, NY, NZ;
    long long NYNZ;
    double dt, dtx, dty, dtz, one_d_dx, one_d_dy, one_d_dz;
};

// Device pointers to the 1-D columns (always FP64: they are tiny and read through the scalar cache).
struct Cols {
    const double *zu, *zw, *zsu, *zsw, *mu, *mw, *rhou, *rhow, *one_d_rhou, *one_d_rhow;
    const double *tb, *tbv, *tbw, *qb, *pib, *f;
};

// Mixed-radix plan of one transform length (radix 4/2/3/5 passes).
constexpr int kMaxRad = 8;
struct Radix {
    int r[kMaxRad];
    int n, nr;
};

// Pressure-solver plan (serial periodic branch: ni = NX-6, nj = NY, levels 1..NZ-2).
struct PressurePlan {
    int ni = 0, nj = 0, nk2 = 0;      // transform sizes and active level count
    int radx[8] = {0}, nradx = 0;      // radix sequence along i
    int rady[8] = {0}, nrady = 0;      // radix sequence along j
    double *twx = nullptr, *twy = nullptr;   // exp(-2 pi i e / n), e = 0..n-1 (interleaved re, im)
    double *coskx = nullptr, *cosly = nullptr; // cos(kwave[i]), cos(lwave[j]) as init_fft3d evaluates them
    double *acoef = nullptr, *ccoef = nullptr; // pressure.cpp:779-782
    void *spec = nullptr;              // spectral work array [kx][ky][kk] complex (T)
    double *ffac = nullptr;            // Thomas multipliers F2 [kx][kk][ky] (FP64)
    double *zpiv = nullptr;            // Thomas pivot reciprocals [kx][kk][ky] (FP64)
    // transform lengths whose factors are all 2, 3 and 5 (the in-place LDS transforms' butterflies); a
    // length with a larger prime factor takes the out-of-place transforms, whose passes sum any radix
    bool smooth_x = true, smooth_y = true;
    bool fused_y = false;              // j-transforms and the vertical solve in one LDS-resident kernel
    bool half_y = false;               // ... or in two per kx on half slabs (even nj; one radix-2 split)
    void *spec2 = nullptr;             // output of the half-slab j pass ([kx][j][kk] complex, T)
    size_t lds_max = 0;
    int cus = 0;  // compute units of the device
};

// per-kernel HIP-event timing on the context stream (eager launches only)
struct TimerRec {
    const char *name;
    hipEvent_t a, b;
    double units;  // gridpoints the launch processes
};

// One member of an ensemble as the batched kernels see it: a device-resident table entry, fixed for the
// ensemble's lifetime.  The three buffers of every prognostic are listed in creation order; which one
// plays M, current or forecast is passed to each launch as EnsRoles (all members rotate in step).
struct EnsMember {
    Geo g;
    Cols c;
    void *buf[7][3];
    const void *bar[6];
    void *pi, *vt;
    const double *sponge_n, *sponge_s;
    int js, jn;
    // pressure plan
    void *spec;
    const double *twx, *twy, *ffac, *zpiv, *coskx, *cosly, *acoef;
    Radix rx, ry;
    int ni, nj, nk2, kcx;
    double dx, dz;
    // first work item of this member in each batched launch
    int mom0, scal0, post0, px0, py0, cp0;
    int nbj_mom, nbj_scal;
};
struct EnsRoles {
    unsigned char r[7][3];  // buffer index of (field, role)
};
// member owning work item `item` of a batched launch whose per-member offsets are the field `first`
// (offsets are non-decreasing; a member without work has the next member's offset)
__device__ __forceinline__ int ens_member(const EnsMember *mem, int nmem, int item, int EnsMember::*first) {
    int lo = 0, hi = nmem - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) / 2;
        if (mem[mid].*first <= item) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}
// Blocks are dispatched round-robin over the 8 XCDs (block b runs on XCD b % 8): tile index of block b
// of n such that each XCD takes one contiguous run of tiles, so neighbouring tiles share that XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int n) {
    const int q = n / 8, rem = n % 8, x = b % 8;
    return x * q + min(x, rem) + b / 8;
}
// the two halves of a fused stage (launch_fused_advect's `parts`)
constexpr int kStageMomentum = 1, kStageScalars = 2;

struct AdjState;  // trajectory and adjoint buffers of the hand-written adjoint (dyn_adjoint.hip)

// output staging (dyn_output.hip): the output fields narrowed to float in one device buffer, copied to
// pinned host memory on a stream of their own
struct OutputStage {
    float *dev = nullptr, *host = nullptr;
    size_t values = 0;
    hipStream_t copy = nullptr;
    hipEvent_t snap = nullptr, done = nullptr;
    bool pending = false;
};
int output_release(struct Ctx *c);

// turbulence parameterization (dyn_turbulence.hip): buf holds Kmix, KHmix, the u/v/w friction and the
// theta/qv/qc/qr diffusion tendencies (storage type, zero outside the interior they are computed on);
// sat is Kessler's saturation flag per point
struct Turbulence {
    bool on = false;
    void *buf[9] = {nullptr};
    unsigned char *sat = nullptr;
    double *landsea = nullptr, *vml2 = nullptr, *vkmax = nullptr;
    double kmixh_max = 0, lH = 0;
};

struct Ctx {
    dyn_grid g{};
    OutputStage out{};
    Turbulence turb{};
    int device = 0;       // the HIP device current at dyn_create: every launch of the context goes there
    Geo geo{};
    int prec = DYN_FP64;
    size_t elt = 8;       // bytes per field element
    size_t npts = 0;      // NX*NY*NZ
    hipStream_t stream = nullptr;
    // second stream of the fused step: the scalar half of a stage (advection + boundary update) runs on
    // it beside the momentum half and the pressure solve, forked and joined by the two events
    hipStream_t side = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // prognostics: [field][role] device pointers; roles rotate by pointer swaps
    void *prog[7][3] = {{nullptr}};
    void *pi = nullptr, *vt = nullptr, *accrain = nullptr;
    void *bar[6] = {nullptr};  // ubar, vbar, wbar, thbar, qbar, pbar
    double *colbuf = nullptr;  // all columns, device
    double *colhost = nullptr; // host copy
    Cols cols{};
    bool cols_set[DYN_COLUMN_COUNT] = {false};
    double *sponge_n = nullptr, *sponge_s = nullptr; // device sponge coefficients
    double *raycoef = nullptr;                       // device Rayleigh coefficients [2][NZ]
    int ray_level_cached = -1;
    int *dflag = nullptr;                             // device error flag (Kessler guard)
    int *fflag = nullptr;                             // device bit mask of fields holding a NaN or Inf
    PressurePlan pp;
    // graph for dyn_step (per-operator steps: two steps restore the buffer assignment)
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    // Between the fused steps of one dyn_step / dyn_step_fused call the M and current roles of the
    // stepped prognostics share a buffer (both hold the state at the start of the step); spare[f] is
    // then the field's third buffer, nullptr otherwise.  Calls end with M and current apart again.
    void *spare[7] = {nullptr};
    // graphs of 6, 3 or 2 fused steps that restore the buffer assignment they start from, keyed by
    // that assignment
    struct StepGraph {
        int steps;
        void *key[7][3];
        hipGraph_t graph;
        hipGraphExec_t exec;
    };
    std::vector<StepGraph> graphs;
    AdjState *adj = nullptr;
    // launch-shape knobs (DYN_TUNE_*), 0 = default
    int tune[DYN_TUNE_COUNT] = {0};
    // timing
    bool timing = false;
    bool capturing = false;
    std::vector<TimerRec> recs;
    std::vector<hipEvent_t> pool;
};

// Records a start event at construction and a stop event at destruction when timing is on.
class Timed {
  public:
    Timed(Ctx *c, const char *name, double units);
    ~Timed();

  private:
    Ctx *c_;
    long idx_ = -1;
};

inline size_t field_count(const Ctx *c, int field) {
    return field == DYN_ACCRAIN ? (size_t)c->g.nx * c->g.ny : c->npts;
}

// launchers implemented in the .hip files (return DYN_* codes)
int launch_advect_uvw(Ctx *c, double step, int il, int ih, int jl, int jh);
int launch_advect_theta(Ctx *c, double step, int il, int ih, int jl, int jh);
int launch_advect_moisture(Ctx *c, double step, int il, int ih, int jl, int jh);
int launch_zero_moisture(Ctx *c, int il, int ih, int jl, int jh);
int launch_pgf(Ctx *c, double step, int il, int ih, int jl, int jh);
int launch_bc(Ctx *c, bool micro);
int launch_periodic(Ctx *c, void *const *fields, int nfields);
int launch_kessler(Ctx *c, int il, int ih, int jl, int jh);
int launch_kessler_step_end(Ctx *c, int ray_level);
// turbulence: tendencies at the start of a step, and their addition to the forecasts in a stage
void turb_release(Ctx *c);
int turb_step_begin(Ctx *c);
int turb_apply_velocity(Ctx *c, double step);
int turb_apply_scalars(Ctx *c, double step, bool theta);
void drop_graph(Ctx *c);
int launch_rayleigh(Ctx *c, int il, int ih, int jl, int jh, int level);
int rayleigh_table(Ctx *c, int level);
int pressure_init(Ctx *c);
int pressure_free(Ctx *c);
// after_i (optional) is called with arg once the first pass (divergence and forward i transform) is
// queued; a nonzero return ends the solve with that code
using PressureHook = int (*)(Ctx *c, void *arg);
int launch_solve_pressure(Ctx *c, double step, bool halos = true, PressureHook after_i = nullptr, void *arg = nullptr);
int launch_pressure_adjoint(Ctx *c, double step, void *pihat, void *uh, void *vh, void *wh, void *r1, void *r2);
int launch_convert(Ctx *c, void *dst, const double *src, size_t n, bool to_device_type);
void adjoint_release(Ctx *c);
// a whole RK3 step with a callback at the start of each stage (stage 0, 1, 2); a non-zero return aborts
typedef int (*StageHook)(Ctx *c, int stage, void *arg);
// share_end: let the fused step end with M and current sharing a buffer (the next step must follow
// through step_with_hook too, and the last one of a sequence must not share)
int step_with_hook(Ctx *c, StageHook hook, void *arg, bool share_end = false);
void unshare_roles(Ctx *c);
bool fused_supported(const Ctx *c);
int fused_waves(const Ctx *c);
// ensembles: batched launches over a device table of members (dyn_fused.hip, dyn_pressure.hip)
int ens_prepare(bool fp32);
int launch_ens_stage(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, const int *shape, int tot_mom, int tot_scal);
int launch_ens_post(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, bool pgf, int total);
int launch_ens_copy(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, int total);
bool pressure_ens_fill(Ctx *c, EnsMember &e, size_t &lds_x, size_t &lds_y);
int launch_pressure_ens(Ctx *c0, const EnsMember *d, int nmem, const EnsRoles &ro, double step, int tot_px, int tot_py, size_t lds_x,
                        size_t lds_y);
// launch shapes of the stage kernels (DYN_TUNE_*): validity and the value in force
bool fused_shape_ok(const Ctx *c, int param, int value);
int fused_shape(const Ctx *c, int param);

}  // namespace dyn
