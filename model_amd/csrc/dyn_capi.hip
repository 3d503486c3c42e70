// C ABI of libdyncore (include/dyncore.h): context lifetime, HBM-resident state, transfers, the
// per-operator entry points and whole-step drivers.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "dyn_internal.hpp"

namespace dyn {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

// Device staging of dyn_get_field_noalloc: an FP32 field is widened to double here before its copy to
// the host, so the call allocates nothing.  One array per GPU (a file-scope __device__ array), held until
// the process ends; every call on a GPU shares it, so calls that overlap on one GPU race.  Nothing needs
// clearing: each call overwrites the values it copies out.
constexpr size_t kNoallocScratchDoubles = size_t(8) << 20;  // 64 MiB: a field of up to 8,388,608 values
static __device__ __align__(256) double scratch[kNoallocScratchDoubles];
constexpr int kMaxDevices = 64;
static double *noalloc_scratch_ptr[kMaxDevices] = {};
static std::mutex noalloc_scratch_mutex;

// once per device, with that device current (dyn_create calls it): the array's address on the device.
// In host code `scratch` names a host-side stand-in, never usable as a device address.
static int noalloc_scratch_setup(int dev) {
    if (dev < 0 || dev >= kMaxDevices) {
        set_error("dyn_create: device index beyond the staging table");
        return DYN_E_ARG;
    }
    std::lock_guard<std::mutex> lock(noalloc_scratch_mutex);
    if (noalloc_scratch_ptr[dev]) return DYN_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) {
        set_error("dyn_create: hipGetSymbolAddress of the staging array failed");
        return DYN_E_HIP;
    }
    noalloc_scratch_ptr[dev] = static_cast<double *>(p);
    return DYN_OK;
}

static hipEvent_t take_event(Ctx *c) {
    if (!c->pool.empty()) {
        hipEvent_t e = c->pool.back();
        c->pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

Timed::Timed(Ctx *c, const char *name, double units) : c_(c) {
    if (!c->timing || c->capturing) return;
    TimerRec r{name, take_event(c), take_event(c), units};
    if (!r.a || !r.b) return;
    (void)hipEventRecord(r.a, c->stream);
    idx_ = (long)c->recs.size();
    c->recs.push_back(r);
}

Timed::~Timed() {
    if (idx_ >= 0) (void)hipEventRecord(c_->recs[idx_].b, c_->stream);
}

int pressure_coefficients(Ctx *c);
bool fused_supported(const Ctx *c);
int fused_prepare(Ctx *c);
int launch_fused_advect(Ctx *c, double step, int il, int ih, int jl, int jh, int parts);
int launch_zero_moisture_shell(Ctx *c, int il, int ih, int jl, int jh);
// interior_done: the step-end update of the columns it changes only at the lid is done (by Kessler)
int launch_post_stage(Ctx *c, double step, bool pgf, bool scalars, bool clip, bool uvw, int ray_level = -1,
                      bool interior_done = false);

// the captured step graph bakes in launch shapes and the kernel sequence: drop it when a knob or the
// turbulence switch changes
void drop_graph(Ctx *c) {
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    if (c->graph) (void)hipGraphDestroy(c->graph);
    c->graph_exec = nullptr;
    c->graph = nullptr;
    for (auto &sg : c->graphs) {
        (void)hipGraphExecDestroy(sg.exec);
        (void)hipGraphDestroy(sg.graph);
    }
    c->graphs.clear();
}

namespace {



__global__ void k_copy(double *__restrict__ dst, const double *__restrict__ src, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long s = (long long)gridDim.x * blockDim.x;
    for (; p < n; p += s) dst[p] = src[p];
}
__global__ void k_copy_f(float *__restrict__ dst, const float *__restrict__ src, long long n) {
    long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long s = (long long)gridDim.x * blockDim.x;
    for (; p < n; p += s) dst[p] = src[p];
}
// advance_outer's M <- P copies of several fields in one launch (grid.y = field), 16 bytes per lane and
// 4-byte words for the tail; `words` = bytes per field / 4
struct FieldCopies {
    const void *src[7];
    void *dst[7];
};
__global__ void __launch_bounds__(256) k_copy_fields(FieldCopies fc, long long words) {
    const uint4 *s = (const uint4 *)fc.src[blockIdx.y];
    uint4 *d = (uint4 *)fc.dst[blockIdx.y];
    const long long n16 = words / 4;
    for (long long q = (long long)blockIdx.x * blockDim.x + threadIdx.x; q < n16; q += (long long)gridDim.x * blockDim.x) d[q] = s[q];
    if (blockIdx.x == 0 && threadIdx.x < (unsigned)(words - n16 * 4))
        ((unsigned *)d)[n16 * 4 + threadIdx.x] = ((const unsigned *)s)[n16 * 4 + threadIdx.x];
}
__global__ void k_clear_flag(int *f) { *f = 0; }

// The frame of a field: whole columns with i < 6, i >= NX-6, j < 6 or j >= NY-6 (NX, NY >= 12).  The
// twelve outer i-slabs are contiguous runs of NY*NZ points, the rest two runs of 6*NZ per i.
template <typename T>
__global__ void __launch_bounds__(256) k_copy_frame(FieldCopies fc, int NX, int NY, int NZ) {
    const T *s = (const T *)fc.src[blockIdx.y];
    T *d = (T *)fc.dst[blockIdx.y];
    const long long slab = (long long)NY * NZ, na = 12 * slab, run = 12LL * NZ, nb = (long long)(NX - 12) * run;
    for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < na + nb; e += (long long)gridDim.x * blockDim.x) {
        long long p;
        if (e < na) {
            const int sl = (int)(e / slab);
            p = (long long)(sl < 6 ? sl : NX - 12 + sl) * slab + e % slab;
        } else {
            const long long q = e - na, r = q % run;
            const int i = 6 + (int)(q / run), jj = (int)(r / NZ);
            p = ((long long)i * NY + (jj < 6 ? jj : NY - 12 + jj)) * NZ + r % NZ;
        }
        d[p] = s[p];
    }
}

int device_copy(Ctx *c, void *dst, const void *src) {
    long long n = (long long)c->npts;
    unsigned blocks = (unsigned)std::min<long long>((n + 255) / 256, 8192);
    Timed _t(c, "advance_copy", (double)n);
    if (c->prec == DYN_FP32) hipLaunchKernelGGL(k_copy_f, dim3(blocks), dim3(256), 0, c->stream, (float *)dst, (const float *)src, n);
    else hipLaunchKernelGGL(k_copy, dim3(blocks), dim3(256), 0, c->stream, (double *)dst, (const double *)src, n);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

// M <- P of fields [f0, f1] in one launch (hipMalloc'd buffers: 256-byte aligned)
int device_copy_fields(Ctx *c, int f0, int f1) {
    FieldCopies fc{};
    const int nf = f1 - f0 + 1;
    for (int f = f0; f <= f1; f++) {
        fc.src[f - f0] = c->prog[f][DYN_ROLE_P];
        fc.dst[f - f0] = c->prog[f][DYN_ROLE_M];
    }
    const long long words = (long long)(c->npts * c->elt / 4);
    const unsigned blocks = (unsigned)std::max<long long>(1, std::min<long long>((words / 4 + 255) / 256, 2048));
    Timed _t(c, "advance_copy", (double)c->npts * nf);
    hipLaunchKernelGGL(k_copy_fields, dim3(blocks, nf), dim3(256), 0, c->stream, fc, words);
    DYN_LAUNCH_CHECK();
    return DYN_OK;
}

// Stage 2 of a step that starts with M and current sharing a buffer writes the spare buffer, which then
// holds an older state than the reference's buffer (the step's starting state).  Stage 2 reads some
// points of that buffer before it writes them: w's levels 0, 1 and NZ-1 everywhere (the divergence
// reads w there) and, of every field, points of the halo and of the columns next to it (the
// divergence's halo columns; the periodic copy carries levels 0 and NZ-1 of the source columns before
// the lid).  The reference's step sequence, run on the CPU with the rest of that buffer poisoned,
// gives the same bits as without the poison, so copying w whole and the six-column frame of the other
// fields from the starting state gives every value the reference has.
// to_spare: copy into the spare buffers (stage 2's output buffers while stage 1 still runs) instead of
// the forecast role.
int device_copy_shell(Ctx *c, int f0, int f1, bool to_spare = false) {
    const int NX = c->g.nx, NY = c->g.ny, NZ = c->g.nz;
    const long long words = (long long)(c->npts * c->elt / 4);
    const unsigned wblocks = (unsigned)std::max<long long>(1, std::min<long long>((words / 4 + 255) / 256, 2048));
    Timed _t(c, "shell_copy", (double)c->npts);
    FieldCopies whole{}, frame{};
    int nw = 0, nf = 0;
    for (int f = f0; f <= f1; f++) {
        void *dst = to_spare ? c->spare[f] : c->prog[f][DYN_ROLE_P];
        if (!dst) {
            set_error("fused step: no spare buffer for the stage-2 copy");
            return DYN_E_STATE;
        }
        const bool all = f == DYN_W || NX < 12 || NY < 12;  // narrow grids: no interior beyond the frame
        (all ? whole.src[nw] : frame.src[nf]) = c->prog[f][DYN_ROLE_M];
        (all ? whole.dst[nw++] : frame.dst[nf++]) = dst;
    }
    hipLaunchKernelGGL(k_copy_fields, dim3(wblocks, nw), dim3(256), 0, c->stream, whole, words);
    DYN_LAUNCH_CHECK();
    if (nf > 0) {
        const long long n = 12LL * NY * NZ + (long long)(NX - 12) * 12 * NZ;
        const dim3 grid((unsigned)std::min<long long>((n + 255) / 256, 1024), nf);
        if (c->prec == DYN_FP32) hipLaunchKernelGGL(k_copy_frame<float>, grid, dim3(256), 0, c->stream, frame, NX, NY, NZ);
        else hipLaunchKernelGGL(k_copy_frame<double>, grid, dim3(256), 0, c->stream, frame, NX, NY, NZ);
        DYN_LAUNCH_CHECK();
    }
    return DYN_OK;
}

Ctx *as_ctx(dyn_ctx *p) { return reinterpret_cast<Ctx *>(p); }

bool cols_ready(const Ctx *c) {
    static const int need[] = {DYN_COL_ZU, DYN_COL_ZW, DYN_COL_MU, DYN_COL_MW, DYN_COL_RHOU, DYN_COL_RHOW, DYN_COL_ONE_D_RHOU,
                               DYN_COL_TB, DYN_COL_TBV, DYN_COL_QB, DYN_COL_F};
    for (int k : need)
        if (!c->cols_set[k]) return false;
    return true;
}

#define CHECK_CTX(c)                                              \
    do {                                                          \
        if (!(c)) {                                               \
            set_error("null context");                            \
            return DYN_E_ARG;                                     \
        }                                                         \
    } while (0)
#define CHECK_COLS(c)                                                                                        \
    do {                                                                                                     \
        if (!cols_ready(c)) {                                                                                \
            set_error("columns (zu, zw, mu, mw, rhou, rhow, one_d_rhou, tb, tbv, qb, f) must be set first"); \
            return DYN_E_STATE;                                                                              \
        }                                                                                                    \
    } while (0)

int check_box(const Ctx *c, int il, int ih, int jl, int jh, int lo_i, int hi_i, int lo_j, int hi_j, const char *what) {
    if (il < lo_i || ih > hi_i || jl < lo_j || jh > hi_j || il >= ih || jl >= jh) {
        set_error(std::string(what) + ": bounds [" + std::to_string(il) + "," + std::to_string(ih) + ")x[" + std::to_string(jl) +
                  "," + std::to_string(jh) + ") outside the stencil-safe range [" + std::to_string(lo_i) + "," +
                  std::to_string(hi_i) + ")x[" + std::to_string(lo_j) + "," + std::to_string(hi_j) + ")");
        return DYN_E_ARG;
    }
    return DYN_OK;
}

int upload_columns(Ctx *c) {
    size_t bytes = sizeof(double) * ((size_t)(DYN_COLUMN_COUNT - 1) * c->g.nz + c->g.ny);
    DYN_HIP(hipMemcpyAsync(c->colbuf, c->colhost, bytes, hipMemcpyHostToDevice, c->stream));
    return DYN_OK;
}

int swap_roles(Ctx *c, int f0, int f1) {
    for (int f = f0; f <= f1; f++) std::swap(c->prog[f][DYN_ROLE_CUR], c->prog[f][DYN_ROLE_P]);
    return DYN_OK;
}

// NaN / Inf scan: one bit per field in *flag; one atomic per wavefront that saw a non-finite value
template <typename T>
__global__ void k_nonfinite(const T *__restrict__ a, long long n, int *__restrict__ flag, int bit) {
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        bad |= !isfinite(a[i]);
    if (__any(bad) && (threadIdx.x % warpSize) == 0) atomicOr(flag, bit);
}

}  // namespace
}  // namespace dyn

using namespace dyn;

extern "C" {

const char *dyn_last_error(void) { return g_last_error.c_str(); }
int dyn_version(void) { return 1; }

int dyn_create(const dyn_grid *grid, dyn_ctx **out) {
    if (!grid || !out) {
        set_error("dyn_create: null argument");
        return DYN_E_ARG;
    }
    *out = nullptr;
    const dyn_grid &g = *grid;
    if (g.nx < 12 || g.ny < 8 || g.nz < 5 || !(g.dx > 0) || !(g.dz > 0) || !(g.dt > 0)) {
        set_error("dyn_create: grid too small or non-positive spacing");
        return DYN_E_ARG;
    }
    if (g.periodic_ew != 1) {
        set_error("dyn_create: only the serial periodic east-west path (PERIODIC_BOUNDARIES=1) is implemented");
        return DYN_E_UNSUPPORTED;
    }
    if (g.microphysics != 0 && g.microphysics != 1) {
        set_error("dyn_create: microphysics option must be 0 (dry) or 1 (Kessler)");
        return DYN_E_UNSUPPORTED;
    }
    if (g.precision != DYN_FP64 && g.precision != DYN_FP32) {
        set_error("dyn_create: precision must be DYN_FP64 or DYN_FP32");
        return DYN_E_ARG;
    }
    if (g.sponge_north < 0 || g.sponge_south < 0 || g.sponge_north + g.sponge_south > g.ny) {
        set_error("dyn_create: sponge widths out of range");
        return DYN_E_ARG;
    }
    if (g.rayleigh_level == g.nz - 1) {
        set_error("dyn_create: rayleigh_level = NZ-1 gives 0/0 in damping.cpp:1159-1160; use <= NZ-2 or >= NZ");
        return DYN_E_ARG;
    }
    if (g.rayleigh_level < 1) {
        set_error("dyn_create: rayleigh_level must be >= 1");
        return DYN_E_ARG;
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) {
        set_error("dyn_create: no HIP device");
        return DYN_E_HIP;
    }
    if (int rc = noalloc_scratch_setup(dev)) return rc;
    Ctx *c = new Ctx();
    c->g = g;
    c->device = dev;
    c->prec = g.precision;
    c->elt = g.precision == DYN_FP32 ? 4 : 8;
    c->npts = (size_t)g.nx * g.ny * g.nz;
    Geo &geo = c->geo;
    geo.NX = g.nx;
    geo.NY = g.ny;
    geo.NZ = g.nz;
    geo.NYNZ = (long long)g.ny * g.nz;
    geo.dt = g.dt;
    geo.dtx = g.dt / g.dx;  // initializer.cpp:173-179, dy = dx
    geo.dty = g.dt / g.dx;
    geo.dtz = g.dt / g.dz;
    geo.one_d_dx = 1.0 / g.dx;
    geo.one_d_dy = 1.0 / g.dx;
    geo.one_d_dz = 1.0 / g.dz;
    auto fail = [&](int rc) {
        dyn_destroy(reinterpret_cast<dyn_ctx *>(c));
        return rc;
    };
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        delete c;
        return DYN_E_HIP;
    }
    if (hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming) != hipSuccess) {
        set_error("second stream or its events could not be created");
        return fail(DYN_E_HIP);
    }
    size_t bytes = c->npts * c->elt;
    int nprog = g.microphysics ? 7 : 4;
    for (int f = 0; f < nprog; f++)
        for (int r = 0; r < 3; r++) {
            if (hipMalloc(&c->prog[f][r], bytes) != hipSuccess || hipMemset(c->prog[f][r], 0, bytes) != hipSuccess) {
                set_error("hipMalloc of a prognostic field failed");
                return fail(DYN_E_HIP);
            }
        }
    void **singles[] = {&c->pi, &c->vt, &c->bar[0], &c->bar[1], &c->bar[2], &c->bar[3], &c->bar[4], &c->bar[5]};
    for (void **p : singles) {
        if (hipMalloc(p, bytes) != hipSuccess || hipMemset(*p, 0, bytes) != hipSuccess) {
            set_error("hipMalloc of a diagnostic field failed");
            return fail(DYN_E_HIP);
        }
    }
    size_t acc = (size_t)g.nx * g.ny * c->elt;
    if (hipMalloc(&c->accrain, acc) != hipSuccess || hipMemset(c->accrain, 0, acc) != hipSuccess) return fail(DYN_E_HIP);
    size_t ncol = (size_t)(DYN_COLUMN_COUNT - 1) * g.nz + g.ny;
    c->colhost = (double *)calloc(ncol, sizeof(double));
    if (hipMalloc(&c->colbuf, ncol * sizeof(double)) != hipSuccess || hipMemset(c->colbuf, 0, ncol * sizeof(double)) != hipSuccess)
        return fail(DYN_E_HIP);
    const double *cb = c->colbuf;
    const double **cp[] = {&c->cols.zu, &c->cols.zw, &c->cols.zsu, &c->cols.zsw, &c->cols.mu, &c->cols.mw, &c->cols.rhou, &c->cols.rhow,
                           &c->cols.one_d_rhou, &c->cols.one_d_rhow, &c->cols.tb, &c->cols.tbv, &c->cols.tbw, &c->cols.qb, &c->cols.pib,
                           &c->cols.f};
    for (int k = 0; k < DYN_COLUMN_COUNT; k++) *cp[k] = cb + (size_t)k * g.nz;
    // sponge coefficients: init_boundaries via cos_profile (boundaries.cpp:84-86, util.cpp:740-749)
    auto cos_profile = [](double pl, double ph, double frac, double shift) {
        if (frac >= 0 && frac <= 1.0) return std::cos(pl * 2 * kTrigPi + (ph - pl) * 2 * kTrigPi * frac) + shift;
        return 0.0;
    };
    std::vector<double> sn(std::max(1, g.sponge_north)), ss(std::max(1, g.sponge_south));
    for (int i = 0; i < g.sponge_north; i++) sn[i] = 0.5 * cos_profile(-0.5, 0, (double)(i + 1) / (double)g.sponge_north, 1);
    for (int i = 0; i < g.sponge_south; i++) ss[i] = 0.5 * cos_profile(0, 0.5, (double)i / (double)g.sponge_south, 1);
    if (hipMalloc(&c->sponge_n, sn.size() * 8) != hipSuccess || hipMalloc(&c->sponge_s, ss.size() * 8) != hipSuccess ||
        hipMemcpy(c->sponge_n, sn.data(), sn.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(c->sponge_s, ss.data(), ss.size() * 8, hipMemcpyHostToDevice) != hipSuccess)
        return fail(DYN_E_HIP);
    if (hipMalloc(&c->raycoef, sizeof(double) * 2 * g.nz) != hipSuccess || hipMemset(c->raycoef, 0, sizeof(double) * 2 * g.nz) != hipSuccess)
        return fail(DYN_E_HIP);
    if (hipMalloc(&c->dflag, sizeof(int)) != hipSuccess || hipMemset(c->dflag, 0, sizeof(int)) != hipSuccess) return fail(DYN_E_HIP);
    if (hipMalloc(&c->fflag, sizeof(int)) != hipSuccess) return fail(DYN_E_HIP);
    int rc = pressure_init(c);
    if (rc) return fail(rc);
    if (fused_supported(c)) {
        rc = fused_prepare(c);
        if (rc) return fail(rc);
    }
    if (hipDeviceSynchronize() != hipSuccess) return fail(DYN_E_HIP);
    *out = reinterpret_cast<dyn_ctx *>(c);
    return DYN_OK;
}

int dyn_destroy(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    if (!c) return DYN_OK;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (auto &r : c->recs) {
        (void)hipEventDestroy(r.a);
        (void)hipEventDestroy(r.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    drop_graph(c);
    unshare_roles(c);
    adjoint_release(c);
    output_release(c);
    turb_release(c);
    pressure_free(c);
    for (auto &f : c->prog)
        for (void *p : f)
            if (p) (void)hipFree(p);
    void *ptrs[] = {c->pi, c->vt, c->accrain, c->bar[0], c->bar[1], c->bar[2], c->bar[3], c->bar[4], c->bar[5], c->colbuf,
                    c->sponge_n, c->sponge_s, c->raycoef, c->dflag, c->fflag};
    for (void *p : ptrs)
        if (p) (void)hipFree(p);
    free(c->colhost);
    if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
    if (c->ev_join) (void)hipEventDestroy(c->ev_join);
    if (c->side) (void)hipStreamDestroy(c->side);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
    return DYN_OK;
}

int dyn_set_column(dyn_ctx *ctx, int column, const double *values, size_t count) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (column < 0 || column >= DYN_COLUMN_COUNT || !values) {
        set_error("dyn_set_column: bad column id");
        return DYN_E_ARG;
    }
    size_t need = column == DYN_COL_F ? (size_t)c->g.ny : (size_t)c->g.nz;
    if (count != need) {
        set_error("dyn_set_column: expected " + std::to_string(need) + " values");
        return DYN_E_ARG;
    }
    memcpy(c->colhost + (size_t)column * c->g.nz, values, need * sizeof(double));
    c->cols_set[column] = true;
    int rc = upload_columns(c);
    if (rc) return rc;
    if (column == DYN_COL_MU || column == DYN_COL_MW || column == DYN_COL_RHOW || column == DYN_COL_RHOU) {
        if (c->cols_set[DYN_COL_MU] && c->cols_set[DYN_COL_MW] && c->cols_set[DYN_COL_RHOW] && c->cols_set[DYN_COL_RHOU]) {
            DYN_HIP(hipStreamSynchronize(c->stream));
            rc = pressure_coefficients(c);
            if (rc) return rc;
        }
    }
    c->ray_level_cached = -1;
    return DYN_OK;
}

static void *field_ptr(Ctx *c, int field, int role) {
    if (field >= DYN_U && field <= DYN_QR) {
        if (role < 0 || role > 2) return nullptr;
        return c->prog[field][role];
    }
    switch (field) {
    case DYN_PI: return c->pi;
    case DYN_UBAR: return c->bar[0];
    case DYN_VBAR: return c->bar[1];
    case DYN_WBAR: return c->bar[2];
    case DYN_THBAR: return c->bar[3];
    case DYN_QBAR: return c->bar[4];
    case DYN_PBAR: return c->bar[5];
    case DYN_VT: return c->vt;
    case DYN_ACCRAIN: return c->accrain;
    }
    return nullptr;
}

static int transfer(Ctx *c, int field, int role, double *host, size_t count, bool up) {
    if (field < 0 || field >= DYN_FIELD_COUNT || !host) {
        set_error("bad field id");
        return DYN_E_ARG;
    }
    void *d = field_ptr(c, field, role);
    if (!d) {
        set_error("field not present in this context (moisture fields need microphysics=1) or bad role");
        return DYN_E_ARG;
    }
    size_t n = field_count(c, field);
    if (count != n) {
        set_error("transfer: expected " + std::to_string(n) + " values, got " + std::to_string(count));
        return DYN_E_ARG;
    }
    if (c->prec == DYN_FP64) {
        if (up) DYN_HIP(hipMemcpyAsync(d, host, n * 8, hipMemcpyHostToDevice, c->stream));
        else DYN_HIP(hipMemcpyAsync(host, d, n * 8, hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(hipStreamSynchronize(c->stream));
        return DYN_OK;
    }
    double *tmp = nullptr;
    DYN_HIP(hipMalloc(&tmp, n * 8));
    int rc = DYN_OK;
    if (up) {
        if (hipMemcpyAsync(tmp, host, n * 8, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = DYN_E_HIP;
        if (!rc) rc = launch_convert(c, d, tmp, n, true);
    } else {
        rc = launch_convert(c, tmp, (const double *)d, n, false);
        if (!rc && hipMemcpyAsync(host, tmp, n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = DYN_E_HIP;
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = DYN_E_HIP;
    (void)hipFree(tmp);
    if (rc) set_error("FP32 transfer failed");
    return rc;
}

int dyn_get_field_noalloc(dyn_ctx *ctx, int field, int role, double *values, size_t count) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (field < 0 || field >= DYN_FIELD_COUNT || !values) {
        set_error("dyn_get_field_noalloc: bad field id");
        return DYN_E_ARG;
    }
    void *d = field_ptr(c, field, role);
    if (!d) {
        set_error("dyn_get_field_noalloc: field not present in this context (moisture fields need microphysics=1) or bad role");
        return DYN_E_ARG;
    }
    size_t n = field_count(c, field);
    if (count != n) {
        set_error("dyn_get_field_noalloc: expected " + std::to_string(n) + " values, got " + std::to_string(count));
        return DYN_E_ARG;
    }
    if (c->prec == DYN_FP64) {  // no staging: the field is already double
        DYN_HIP(hipMemcpyAsync(values, d, n * 8, hipMemcpyDeviceToHost, c->stream));
        DYN_HIP(hipStreamSynchronize(c->stream));
        return DYN_OK;
    }
    if (n > kNoallocScratchDoubles) {
        set_error("dyn_get_field_noalloc: " + std::to_string(n) + " values exceed the " +
                  std::to_string(kNoallocScratchDoubles) + "-value device staging array; use dyn_get_field");
        return DYN_E_ARG;
    }
    double *stage = noalloc_scratch_ptr[c->device];  // the array's device address on this GPU
    int rc = launch_convert(c, stage, (const double *)d, n, false);
    if (!rc && hipMemcpyAsync(values, stage, n * 8, hipMemcpyDeviceToHost, c->stream) != hipSuccess) rc = DYN_E_HIP;
    if (hipStreamSynchronize(c->stream) != hipSuccess) rc = DYN_E_HIP;
    if (rc) set_error("dyn_get_field_noalloc: FP32 read-back failed");
    return rc;
}

int dyn_set_field(dyn_ctx *ctx, int field, int role, const double *values, size_t count) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    return transfer(c, field, role, const_cast<double *>(values), count, true);
}

int dyn_get_field(dyn_ctx *ctx, int field, int role, double *values, size_t count) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    return transfer(c, field, role, values, count, false);
}

int dyn_synchronize(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    DYN_HIP(hipStreamSynchronize(c->stream));
    int flag = 0;
    DYN_HIP(hipMemcpy(&flag, c->dflag, sizeof(int), hipMemcpyDeviceToHost));
    if (flag) {
        set_error("Kessler evaporation E > 100 (the reference aborts here, kessler.cpp:83-88)");
        return DYN_E_NUMERIC;
    }
    return DYN_OK;
}

int dyn_check_finite(dyn_ctx *ctx, int *bad_fields) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    hipLaunchKernelGGL(k_clear_flag, dim3(1), dim3(1), 0, c->stream, c->fflag);
    const int last = c->g.microphysics ? DYN_QR : DYN_TH;
    for (int f = DYN_U; f <= DYN_PI; f++) {
        if (f > last && f < DYN_PI) continue;
        const void *d = f == DYN_PI ? (const void *)c->pi : (const void *)c->prog[f][DYN_ROLE_CUR];
        const long long n = (long long)c->npts;
        const int blocks = (int)std::min<long long>((n + 255) / 256, 2048);
        if (c->prec == DYN_FP32)
            hipLaunchKernelGGL(k_nonfinite<float>, dim3(blocks), dim3(256), 0, c->stream, (const float *)d, n, c->fflag, 1 << f);
        else
            hipLaunchKernelGGL(k_nonfinite<double>, dim3(blocks), dim3(256), 0, c->stream, (const double *)d, n, c->fflag, 1 << f);
        DYN_HIP(hipGetLastError());
    }
    int mask = 0;
    DYN_HIP(hipMemcpyAsync(&mask, c->fflag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    DYN_HIP(hipStreamSynchronize(c->stream));
    if (bad_fields) *bad_fields = mask;
    if (mask) {
        static const char *names[] = {"u", "v", "w", "theta", "qv", "qc", "qr", "pi"};
        std::string which;
        for (int f = 0; f <= DYN_PI; f++)
            if (mask & (1 << f)) which += std::string(which.empty() ? "" : ", ") + names[f];
        set_error("non-finite values (NaN or Inf) in the current " + which);
        return DYN_E_NUMERIC;
    }
    return DYN_OK;
}

int dyn_advect_uvw(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    int rc = check_box(c, il, ih, jl, jh, 3, c->g.nx - 3, 3, c->g.ny - 3, "advect_uvw_velocity");
    return rc ? rc : launch_advect_uvw(c, step, il, ih, jl, jh);
}

int dyn_advect_theta(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    int rc = check_box(c, il, ih, jl, jh, 3, c->g.nx - 3, 3, c->g.ny - 3, "advect_theta");
    return rc ? rc : launch_advect_theta(c, step, il, ih, jl, jh);
}

int dyn_advect_moisture(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (!c->g.microphysics) {
        set_error("advect_microphysics_cell: context has no moisture fields");
        return DYN_E_STATE;
    }
    int rc = check_box(c, il, ih, jl, jh, 3, c->g.nx - 3, 3, c->g.ny - 3, "advect_microphysics_cell");
    return rc ? rc : launch_advect_moisture(c, step, il, ih, jl, jh);
}

int dyn_zero_moisture(dyn_ctx *ctx, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (!c->g.microphysics) {
        set_error("zero_moisture: context has no moisture fields");
        return DYN_E_STATE;
    }
    int rc = check_box(c, il, ih, jl, jh, 0, c->g.nx, 0, c->g.ny, "zero_moisture");
    return rc ? rc : launch_zero_moisture(c, il, ih, jl, jh);
}

int dyn_solve_pressure(dyn_ctx *ctx, double step) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (!(step > 0)) {
        set_error("solve_pressure: step must be positive (the divergence is divided by step*dt)");
        return DYN_E_ARG;
    }
    return launch_solve_pressure(c, step);
}

int dyn_pgf(dyn_ctx *ctx, double step, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    int rc = check_box(c, il, ih, jl, jh, 1, c->g.nx, 1, c->g.ny, "pressure gradient");
    return rc ? rc : launch_pgf(c, step, il, ih, jl, jh);
}

int dyn_apply_bc(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    return launch_bc(c, false);
}

int dyn_apply_bc_micro(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (!c->g.microphysics) {
        set_error("apply_boundary_condition_microphysics: context has no moisture fields");
        return DYN_E_STATE;
    }
    return launch_bc(c, true);
}

int dyn_kessler(dyn_ctx *ctx, int il, int ih, int jl, int jh) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (!c->g.microphysics) {
        set_error("run_microphysics: context has no moisture fields");
        return DYN_E_STATE;
    }
    int rc = check_box(c, il, ih, jl, jh, 0, c->g.nx, 0, c->g.ny, "run_microphysics");
    return rc ? rc : launch_kessler(c, il, ih, jl, jh);
}

// Rayleigh coefficients exactly as damping.cpp:1153-1160 forms them (unstretched zu, zw), uploaded
// once per level so that no host copy happens inside a captured step
static int ensure_rayleigh_table(Ctx *c, int level) {
    const int NZ = c->g.nz;
    if (c->ray_level_cached == level) return DYN_OK;
    const double *zu = c->colhost + (size_t)DYN_COL_ZU * NZ, *zw = c->colhost + (size_t)DYN_COL_ZW * NZ;
    std::vector<double> coef(2 * (size_t)NZ, 0.0);
    double raydmpcoef = 0.05, raydmpz = zu[level];
    for (int k = level; k < NZ; k++) {
        coef[k] = raydmpcoef * 0.5 * (1.0 - std::cos(kTrigPi * (zu[k] - raydmpz) / (zu[NZ - 1] - raydmpz)));
        coef[NZ + k] = raydmpcoef * 0.5 * (1.0 - std::cos(kTrigPi * (zw[k] - raydmpz) / (zw[NZ - 1] - raydmpz)));
    }
    DYN_HIP(hipMemcpyAsync(c->raycoef, coef.data(), coef.size() * 8, hipMemcpyHostToDevice, c->stream));
    DYN_HIP(hipStreamSynchronize(c->stream));
    c->ray_level_cached = level;
    return DYN_OK;
}

int dyn_rayleigh(dyn_ctx *ctx, int il, int ih, int jl, int jh, int level) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    int rc = check_box(c, il, ih, jl, jh, 0, c->g.nx, 0, c->g.ny, "rayleigh_damping");
    if (rc) return rc;
    const int NZ = c->g.nz;
    if (level >= NZ) return DYN_OK;  // damping.cpp:1156: the k loop is empty
    if (level < 0 || level == NZ - 1) {
        set_error("rayleigh_damping: level must be in [0, NZ-2] or >= NZ (NZ-1 divides 0 by 0)");
        return DYN_E_ARG;
    }
    int rc2 = ensure_rayleigh_table(c, level);
    if (rc2) return rc2;
    return launch_rayleigh(c, il, ih, jl, jh, level);
}

int dyn_advance_inner(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    return swap_roles(c, DYN_U, DYN_TH);
}

int dyn_advance_outer(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    for (int f = DYN_U; f <= DYN_TH; f++) {
        int rc = device_copy(c, c->prog[f][DYN_ROLE_M], c->prog[f][DYN_ROLE_P]);
        if (rc) return rc;
    }
    return swap_roles(c, DYN_U, DYN_TH);
}

int dyn_micro_advance_inner(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (!c->g.microphysics) return DYN_OK;
    return swap_roles(c, DYN_QV, DYN_QR);
}

int dyn_micro_advance_outer(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (!c->g.microphysics) return DYN_OK;
    for (int f = DYN_QV; f <= DYN_QR; f++) {
        int rc = device_copy(c, c->prog[f][DYN_ROLE_M], c->prog[f][DYN_ROLE_P]);
        if (rc) return rc;
    }
    return swap_roles(c, DYN_QV, DYN_QR);
}

// s_integrate_rk3, driver_serial.cpp:178-286 (no diffusion, terrain or heating; turbulence when
// dyn_set_turbulence turned it on)
static int reference_step(dyn_ctx *ctx, StageHook hook = nullptr, void *arg = nullptr) {
    Ctx *c = as_ctx(ctx);
    const double steps[] = {1. / 3., 0.5, 1.0};
    const int NX = c->g.nx, NY = c->g.ny;
    const bool moist = c->g.microphysics != 0;
    int rc;
#define RC(x)                  \
    do {                       \
        if ((rc = (x))) return rc; \
    } while (0)
    const bool turb = c->turb.on;
    if (turb) RC(turb_step_begin(c));  // calculate_diff_tend(1,NX-1,1,NY-1), driver_serial.cpp:194
    for (int s = 0; s < 3; s++) {
        double step = steps[s];
        if (hook) RC(hook(c, s, arg));
        // integrate_non_hydro(step,0,NX,0,NY), solver.cpp:87-138
        RC(dyn_advect_uvw(ctx, step, 3, NX - 3, 3, NY - 3));
        if (turb) RC(turb_apply_velocity(c, step));  // solver.cpp:104
        RC(dyn_solve_pressure(ctx, step));
        RC(dyn_pgf(ctx, step, 1, NX - 1, 1, NY - 1));
        RC(dyn_advect_theta(ctx, step, 3, NX - 3, 3, NY - 3));
        if (turb) RC(turb_apply_scalars(c, step, true));  // driver_serial.cpp:218
        if (moist) {
            RC(dyn_advect_moisture(ctx, step, 3, NX - 3, 3, NY - 3));
            if (turb) RC(turb_apply_scalars(c, step, false));  // driver_serial.cpp:231
            RC(dyn_zero_moisture(ctx, 1, NX - 1, 1, NY - 1));
        }
        RC(dyn_apply_bc(ctx));
        if (moist) {
            RC(dyn_apply_bc_micro(ctx));
            if (s < 2) RC(dyn_micro_advance_inner(ctx));
        }
        if (s < 2) RC(dyn_advance_inner(ctx));
    }
    if (moist) {
        RC(dyn_kessler(ctx, 1, NX - 1, 1, NY - 1));
        RC(dyn_apply_bc_micro(ctx));
        RC(dyn_micro_advance_outer(ctx));
    }
    RC(dyn_apply_bc(ctx));
    RC(dyn_rayleigh(ctx, 1, NX - 1, 1, NY - 1, c->g.rayleigh_level));
    RC(dyn_advance_outer(ctx));
#undef RC
    return DYN_OK;
}

// The scalar half of a fused stage: theta and moisture advection, then zero_moisture outside the
// advected box and the scalars' boundary update.  With fork set it runs on the context's second stream,
// behind an event that orders it after everything the context's stream holds so far, or, with
// recorded set, behind the fork event the caller recorded earlier.  A failed fork is returned; a
// failed launch is kept in rc, so that the caller still joins the second stream.
struct ScalarHalf {
    double step;
    bool fork;
    bool recorded;
    int rc;
};

static int launch_scalar_half(Ctx *c, void *arg) {
    ScalarHalf &h = *(ScalarHalf *)arg;
    const int NX = c->g.nx, NY = c->g.ny;
    const hipStream_t own = c->stream;
    if (h.fork) {
        if ((!h.recorded && hipEventRecord(c->ev_fork, c->stream) != hipSuccess) ||
            hipStreamWaitEvent(c->side, c->ev_fork, 0) != hipSuccess) {
            set_error("fused step: fork onto the second stream failed");
            return DYN_E_HIP;
        }
        c->stream = c->side;
    }
    h.rc = launch_fused_advect(c, h.step, 3, NX - 3, 3, NY - 3, kStageScalars);
    if (!h.rc) h.rc = launch_post_stage(c, h.step, true, true, true, false);
    c->stream = own;
    return DYN_OK;
}

// The same step with the stage advection fused into one kernel (advect_uvw_velocity, advect_theta,
// advect_microphysics_cell and zero_moisture's interior in one pass).  advect_theta and the moisture
// advection read only the current-stage fields, which the pressure solve and the gradient force never
// write, so they move ahead of solve_pressure without changing any value.
//
// share_end: end the step with the M and current roles on one buffer instead of copying M <- P (the
// next step must then be a fused step too, and the call's last step must not share).  A step that
// starts shared reads M and current from one buffer in stage 1, moves current to the stage-1 forecast
// and the forecast role to the spare buffer, and before stage 2 copies the starting state's points
// outside the core into that buffer (device_copy_shell); every other value equals the reference's.
static int fused_step(dyn_ctx *ctx, StageHook hook = nullptr, void *arg = nullptr, bool share_end = false) {
    Ctx *c = as_ctx(ctx);
    const double steps[] = {1. / 3., 0.5, 1.0};
    const int NX = c->g.nx, NY = c->g.ny;
    const bool moist = c->g.microphysics != 0;
    const int flast = moist ? DYN_QR : DYN_TH;  // the prognostics a step rotates
    const bool shared = c->prog[DYN_U][DYN_ROLE_M] == c->prog[DYN_U][DYN_ROLE_CUR];
    int rc;
#define RC(x)                      \
    do {                           \
        if ((rc = (x))) return rc; \
    } while (0)
    // The scalar half of a stage reads the current buffers only and writes buffers nothing else in the
    // stage reads, so it runs on the second stream beside the pressure solve, which needs only the u,
    // v, w forecasts of the momentum half.  Forked after the momentum kernel (the head of the stage's
    // critical path), its blocks fill the CUs the latency-bound transforms leave idle.  Where the j pass
    // runs on half slabs, whose blocks each take most of a CU's LDS, it is forked after the solve's
    // first pass instead (2.08 -> 1.99 ms per FP64 step at 256x256x64; elsewhere neutral or slower).
    // Value 4 keeps the fork after the momentum kernel but queues (and captures) the scalar launches
    // after the solve's first pass, which measured faster for FP64 on the other grids.
    const int streams = fused_shape(c, DYN_TUNE_STAGE_STREAMS);
    const bool two = streams >= 2, late = streams >= 3;
    // With shared roles on two streams, the copy into stage 2's output buffer runs during stage 1, on
    // the second stream after the scalar half: it reads the step's starting state and writes the spare
    // buffer, neither of which stage 1 writes or reads, and the second stream's chain is the shorter one.
    const bool copy_beside = shared && two;
    for (int s = 0; s < 3; s++) {
        double step = steps[s];
        if (hook) RC(hook(c, s, arg));
        if (s == 1 && shared && !copy_beside) RC(device_copy_shell(c, DYN_U, flast));
        RC(launch_fused_advect(c, step, 3, NX - 3, 3, NY - 3, kStageMomentum));
        ScalarHalf half{step, two, streams == 4, DYN_OK};
        if (half.recorded && hipEventRecord(c->ev_fork, c->stream) != hipSuccess) {
            set_error("fused step: fork onto the second stream failed");
            return DYN_E_HIP;
        }
        if (!late) RC(launch_scalar_half(c, &half));
        // the halo columns periodic_uvw_boundaries writes are read only by the divergence (which reads
        // UP(3) for UP(NX-3)); the boundary update overwrites them before anything else reads them
        int rm = launch_solve_pressure(c, step, false, late ? launch_scalar_half : nullptr, &half);
        // pressure gradient and the boundary update of u, v, w
        if (!rm) rm = launch_post_stage(c, step, true, false, false, true);
        if (s == 0 && copy_beside && !rm && !half.rc) {  // the scalar half is on the second stream by now
            const hipStream_t own = c->stream;
            c->stream = c->side;
            rm = device_copy_shell(c, DYN_U, flast, true);
            c->stream = own;
        }
        // the join comes first, also after a failed launch, so no work is left on the second stream
        if (two && (hipEventRecord(c->ev_join, c->side) != hipSuccess || hipStreamWaitEvent(c->stream, c->ev_join, 0) != hipSuccess)) {
            set_error("fused step: join of the second stream failed");
            return DYN_E_HIP;
        }
        RC(half.rc);
        RC(rm);
        // advance_inner / microphysics_advance_inner; from shared roles the forecast role moves to the
        // spare buffer
        if (s < 2)
            for (int f = DYN_U; f <= flast; f++) {
                if (c->prog[f][DYN_ROLE_M] == c->prog[f][DYN_ROLE_CUR]) {
                    c->prog[f][DYN_ROLE_CUR] = c->prog[f][DYN_ROLE_P];
                    c->prog[f][DYN_ROLE_P] = c->spare[f];
                    c->spare[f] = nullptr;
                } else {
                    std::swap(c->prog[f][DYN_ROLE_CUR], c->prog[f][DYN_ROLE_P]);
                }
            }
    }
    // run_microphysics(1, NX-1, 1, NY-1); apply_boundary_condition_microphysics then
    // apply_boundary_condition (different fields, so one kernel serves both; the moisture pointers swap
    // only afterwards); rayleigh_damping(1, NX-1, 1, NY-1) in the same kernels, since it damps each
    // written value pointwise.  Moist steps leave the columns the boundary update changes only at the
    // lid (and by theta's damping) to Kessler's threads, and the boundary kernel does the rest.
    const int lev = c->g.rayleigh_level;
    if (lev < c->g.nz) {
        if (lev < 0 || lev == c->g.nz - 1) {
            set_error("rayleigh_damping: level must be in [0, NZ-2] or >= NZ (NZ-1 divides 0 by 0)");
            return DYN_E_ARG;
        }
        RC(ensure_rayleigh_table(c, lev));
    }
    const int ray = lev < c->g.nz ? lev : -1;
    if (moist) RC(launch_kessler_step_end(c, ray));
    RC(launch_post_stage(c, 0.0, false, true, false, true, ray, moist));
    // microphysics_advance then advance_outer: Rayleigh damping leaves the moisture fields alone, so
    // the M <- P copies of every prognostic go in one launch after it
    if (share_end) {  // M and current on the forecast buffer, the old M buffer spare
        for (int f = DYN_U; f <= flast; f++) {
            c->spare[f] = c->prog[f][DYN_ROLE_M];
            c->prog[f][DYN_ROLE_M] = c->prog[f][DYN_ROLE_P];
            c->prog[f][DYN_ROLE_P] = c->prog[f][DYN_ROLE_CUR];
            c->prog[f][DYN_ROLE_CUR] = c->prog[f][DYN_ROLE_M];
        }
    } else {
        RC(device_copy_fields(c, DYN_U, flast));
        if (moist) RC(swap_roles(c, DYN_QV, DYN_QR));
        RC(swap_roles(c, DYN_U, DYN_TH));
    }
#undef RC
    return DYN_OK;
}

// the sequence dyn_step records: the fused stages when NZ fits one wavefront, else the per-operator
// kernels
static int graph_step(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    if (fused_supported(c)) return fused_step(ctx);
    return reference_step(ctx);
}

int dyn_step_fused(dyn_ctx *ctx, int nsteps) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (!fused_supported(c)) {
        set_error(c->turb.on ? "dyn_step_fused: the fused stage kernels carry no turbulence (dyn_step runs the closure)"
                             : "dyn_step_fused: the fused stage kernels need NZ <= 64 or 66 <= NZ <= 128, and fields below 4 GiB");
        return DYN_E_UNSUPPORTED;
    }
    if (c->g.rayleigh_level < c->g.nz) {
        int rc = ensure_rayleigh_table(c, c->g.rayleigh_level);
        if (rc) return rc;
    }
    // M and current share a buffer between the steps of the call (no M <- P copy), not after its last
    for (int n = 0; n < nsteps; n++) {
        int rc = fused_step(ctx, nullptr, nullptr, n < nsteps - 1);
        if (rc) {
            unshare_roles(c);
            return rc;
        }
    }
    return DYN_OK;
}

int dyn_step_reference(dyn_ctx *ctx, int nsteps) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    for (int n = 0; n < nsteps; n++) {
        int rc = reference_step(ctx);
        if (rc) return rc;
    }
    return DYN_OK;
}

// Graph-replayed fused steps.  A run of k >= 3 steps, k a multiple of 3, whose first step ends with M
// and current sharing a buffer, whose middle steps share, and whose last step separates the roles again
// with the M <- P copy, maps the buffers (M, current, forecast) = (a, b, c) to (b, a, c) (a shared step
// maps (M = current, forecast, spare) = (X, Y, Z) to (Y, Z, X)); M and current then hold the same
// values, so naming them back (a, b, c) changes nothing.  Two steps with the roles apart also restore
// (a, b, c).  Every call of n >= 2 steps is made of runs of 6, 3 and 2 steps, each one graph of the
// current assignment; a context keeps its assignment from call to call, and single steps (which move
// it) run eagerly.  The three graphs of an assignment are captured together when it is first met.
static Ctx::StepGraph *find_step_graph(Ctx *c, int steps) {
    for (auto &g : c->graphs)
        if (g.steps == steps && !memcmp(g.key, c->prog, sizeof(g.key))) return &g;
    return nullptr;
}

static int capture_step_graph(Ctx *c, int steps) {
    dyn_ctx *ctx = reinterpret_cast<dyn_ctx *>(c);
    const int flast = c->g.microphysics ? DYN_QR : DYN_TH;
    Ctx::StepGraph ng{};
    ng.steps = steps;
    memcpy(ng.key, c->prog, sizeof(ng.key));
    DYN_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    c->capturing = true;
    int rc = DYN_OK;
    for (int n = 0; n < steps && !rc; n++) rc = fused_step(ctx, nullptr, nullptr, steps > 2 && n < steps - 1);
    c->capturing = false;
    hipGraph_t graph = nullptr;
    hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (!rc && steps > 2)
        for (int f = DYN_U; f <= flast; f++) std::swap(c->prog[f][DYN_ROLE_M], c->prog[f][DYN_ROLE_CUR]);
    const bool back = !memcmp(ng.key, c->prog, sizeof(ng.key));
    memcpy(c->prog, ng.key, sizeof(ng.key));
    for (auto &p : c->spare) p = nullptr;
    if (rc || e != hipSuccess || !back) {
        if (graph) (void)hipGraphDestroy(graph);
        if (!rc && e != hipSuccess) set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        if (!rc && e == hipSuccess) set_error("internal: a step graph did not restore the buffer assignment");
        return rc ? rc : (e != hipSuccess ? DYN_E_HIP : DYN_E_STATE);
    }
    ng.graph = graph;
    DYN_HIP(hipGraphInstantiate(&ng.exec, graph, nullptr, nullptr, 0));
    c->graphs.push_back(ng);
    return DYN_OK;
}

static int replay_steps(Ctx *c, int steps, int times) {
    if (times <= 0) return DYN_OK;
    if (!find_step_graph(c, 6)) {  // first meeting of this assignment: its three graphs
        int rc = capture_step_graph(c, 6);
        if (!rc) rc = capture_step_graph(c, 3);
        if (!rc) rc = capture_step_graph(c, 2);
        if (rc) return rc;
    }
    Ctx::StepGraph *sg = find_step_graph(c, steps);
    if (!sg) {
        set_error("internal: no step graph for the current buffer assignment");
        return DYN_E_STATE;
    }
    for (int t = 0; t < times; t++) DYN_HIP(hipGraphLaunch(sg->exec, c->stream));
    return DYN_OK;
}

// Steps replayed from hipGraphs.  Fused steps: the first step ends with M and current sharing a buffer,
// the steps between run from graphs of three such steps, the last step separates the roles again with
// the M <- P copy.  Per-operator steps (NZ the fused kernels do not take) permute the current/forecast
// buffers with period 2, so their graph records two steps and an odd trailing step runs eagerly.
int dyn_step(dyn_ctx *ctx, int nsteps) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    CHECK_COLS(c);
    if (nsteps <= 0) return DYN_OK;
    // the Rayleigh table is uploaded with a host copy, which must not happen inside a capture
    if (c->g.rayleigh_level < c->g.nz) {
        int rc = ensure_rayleigh_table(c, c->g.rayleigh_level);
        if (rc) return rc;
    }
    if (fused_supported(c)) {
        if (nsteps == 1) return fused_step(ctx);
        // n = 6 q + (0 | 2 | 3 | 2+2 | 3+2), with 7 = 3+2+2 for r = 1
        int q = nsteps / 6, r = nsteps % 6, n3 = 0, n2 = 0;
        if (r == 1) {
            q -= 1;
            n3 = 1;
            n2 = 2;
        } else {
            n3 = r == 3 || r == 5;
            n2 = r == 2 || r == 5 ? 1 : (r == 4 ? 2 : 0);
        }
        int rc = replay_steps(c, 6, q);
        if (!rc) rc = replay_steps(c, 3, n3);
        if (!rc) rc = replay_steps(c, 2, n2);
        return rc;
    }
    int pairs = nsteps / 2;
    if (pairs > 0) {
        if (!c->graph_exec) {
            void *saved[7][3];
            memcpy(saved, c->prog, sizeof(saved));
            DYN_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            c->capturing = true;
            int rc = graph_step(ctx);
            if (!rc) rc = graph_step(ctx);
            c->capturing = false;
            hipGraph_t graph = nullptr;
            hipError_t e = hipStreamEndCapture(c->stream, &graph);
            if (rc || e != hipSuccess) {
                if (graph) (void)hipGraphDestroy(graph);
                memcpy(c->prog, saved, sizeof(saved));
                if (!rc) set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
                return rc ? rc : DYN_E_HIP;
            }
            // two steps restore the buffer assignment
            if (memcmp(saved, c->prog, sizeof(saved)) != 0) {
                (void)hipGraphDestroy(graph);
                memcpy(c->prog, saved, sizeof(saved));
                set_error("internal: two steps did not restore the buffer assignment");
                return DYN_E_STATE;
            }
            DYN_HIP(hipGraphInstantiate(&c->graph_exec, graph, nullptr, nullptr, 0));
            c->graph = graph;
        }
        for (int p = 0; p < pairs; p++) DYN_HIP(hipGraphLaunch(c->graph_exec, c->stream));
    }
    if (nsteps % 2) {
        int rc = graph_step(ctx);
        if (rc) return rc;
    }
    return DYN_OK;
}

int dyn_timing(dyn_ctx *ctx, int on) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    DYN_HIP(hipStreamSynchronize(c->stream));
    for (auto &r : c->recs) {
        c->pool.push_back(r.a);
        c->pool.push_back(r.b);
    }
    c->recs.clear();
    c->timing = on != 0;
    return DYN_OK;
}

int dyn_timing_report(dyn_ctx *ctx, char *buf, size_t len) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (!buf || len == 0) {
        set_error("dyn_timing_report: empty buffer");
        return DYN_E_ARG;
    }
    DYN_HIP(hipStreamSynchronize(c->stream));
    struct Agg {
        std::string name;
        double ms = 0, units = 0;
        long n = 0;
    };
    std::vector<Agg> agg;
    for (auto &r : c->recs) {
        float ms = 0;
        DYN_HIP(hipEventElapsedTime(&ms, r.a, r.b));
        Agg *a = nullptr;
        for (auto &x : agg)
            if (x.name == r.name) a = &x;
        if (!a) {
            agg.push_back(Agg{r.name});
            a = &agg.back();
        }
        a->ms += ms;
        a->n += 1;
        a->units += r.units;
    }
    std::string out;
    char line[256];
    for (auto &a : agg) {
        snprintf(line, sizeof(line), "%s %ld %.6f %.0f\n", a.name.c_str(), a.n, a.ms, a.units);
        out += line;
    }
    if (out.size() + 1 > len) {
        set_error("dyn_timing_report: buffer too small");
        return DYN_E_ARG;
    }
    memcpy(buf, out.c_str(), out.size() + 1);
    return DYN_OK;
}

double dyn_algorithmic_bytes_per_point(const dyn_ctx *ctx) {
    const Ctx *c = reinterpret_cast<const Ctx *>(ctx);
    if (!c) return 0.0;
    // field passes per step: 93 dry, 135 moist (one read of each input and one write of each
    // output per operator group, spectral arrays counted as one real-equivalent field)
    double passes = c->g.microphysics ? 135.0 : 93.0;
    return passes * (double)c->elt;
}

void *dyn_stream(dyn_ctx *ctx) {
    Ctx *c = as_ctx(ctx);
    return c ? (void *)c->stream : nullptr;
}

int dyn_tune_set(dyn_ctx *ctx, int param, int value) {
    Ctx *c = as_ctx(ctx);
    CHECK_CTX(c);
    if (param < 0 || param >= DYN_TUNE_COUNT) {
        set_error("dyn_tune_set: unknown parameter " + std::to_string(param));
        return DYN_E_ARG;
    }
    if (value != 0 && !fused_shape_ok(c, param, value)) {
        set_error("dyn_tune_set: value " + std::to_string(value) + " not supported for parameter " + std::to_string(param));
        return DYN_E_ARG;
    }
    c->tune[param] = value;
    drop_graph(c);
    return DYN_OK;
}

int dyn_tune_get(dyn_ctx *ctx, int param) {
    Ctx *c = as_ctx(ctx);
    if (!c || param < 0 || param >= DYN_TUNE_COUNT) return DYN_E_ARG;
    return fused_shape(c, param);
}

}  // extern "C"

namespace dyn {
// one step in the fastest eager form, with a callback at the start of every stage
// the M role back on the spare buffer after a sequence that stopped between fused steps (an error): the
// contents are then those of a failed step, but every role has its own buffer again
void unshare_roles(Ctx *c) {
    for (int f = 0; f < 7; f++)
        if (c->spare[f]) {
            c->prog[f][DYN_ROLE_M] = c->spare[f];
            c->spare[f] = nullptr;
        }
}

int step_with_hook(Ctx *c, StageHook hook, void *arg, bool share_end) {
    dyn_ctx *ctx = reinterpret_cast<dyn_ctx *>(c);
    return fused_supported(c) ? fused_step(ctx, hook, arg, share_end) : reference_step(ctx, hook, arg);
}
}  // namespace dyn

namespace dyn {
// the damping coefficients of rayleigh_damping for `level` (host computation and upload; not in a capture)
int rayleigh_table(Ctx *c, int level) { return ensure_rayleigh_table(c, level); }
}  // namespace dyn
