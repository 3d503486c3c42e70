// Ensembles: members of uneven grid sizes advanced together.  Every pass of the step (stage advection,
// the three pressure passes, the boundary epilogue, advance_outer's copy) is one launch over all members:
// the members' work items are concatenated and each block finds its member in a device-resident member
// table.  A step is captured once into a hipGraph (two steps: the buffer rotation has period 2).
//
// The members can be split into groups of about equal point counts (dyn_ensemble_set_groups): each group
// runs the same batched passes over its own members on its own stream, so the passes of one group (say
// the LDS-bound pressure transforms) overlap those of another (the stencils); the captured graph holds one
// branch per group.
//
// All members must share the precision and NZ and be dry (microphysics 0).  The members stay ordinary
// contexts: their field accessors work between ensemble steps after dyn_ensemble_synchronize.
#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "dyn_internal.hpp"

namespace dyn {

// groups after dyn_ensemble_create: as many as the hardware queues a process gets by default (4), beyond
// which further streams share queues
constexpr int kDefaultGroups = 4;

// members advanced by one stream: their table (offsets local to the group's launches) and work totals
struct EnsGroup {
    std::vector<int> idx;
    std::vector<EnsMember> h;
    EnsMember *d = nullptr;
    int tot_mom = 0, tot_scal = 0, tot_post = 0, tot_px = 0, tot_py = 0, tot_cp = 0;
    hipStream_t stream = nullptr;  // the first group runs on the first member's stream
    hipEvent_t done = nullptr;
};

struct Ensemble {
    std::vector<Ctx *> mem;
    std::vector<EnsMember> h;
    EnsMember *d = nullptr;
    int nmem = 0;
    unsigned char perm[7][3];  // buffer index of (field, role)
    int shape[4] = {8, 16, 8, 16};  // momentum TJ, TI, scalar TJ, TI
    int tot_mom = 0, tot_scal = 0, tot_post = 0, tot_px = 0, tot_py = 0, tot_cp = 0;
    bool batched_pressure = false;
    size_t lds_x = 0, lds_y = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<EnsGroup> groups;
    hipEvent_t fork = nullptr;
    Ctx *c0() const { return mem[0]; }
};

namespace {

EnsRoles roles_of(const Ensemble *e) {
    EnsRoles r;
    memcpy(r.r, e->perm, sizeof(r.r));
    return r;
}

// work offsets of `h` (members launched together) for the current launch shapes
void offsets(const Ensemble *e, std::vector<EnsMember> &h, int &tmom, int &tscal, int &tpost, int &tpx, int &tpy, int &tcp) {
    int mom = 0, scal = 0, post = 0, px = 0, py = 0, cp = 0;
    for (EnsMember &t : h) {
        const int nx = t.g.NX - 6, ny = t.g.NY - 6;
        t.nbj_mom = (ny + e->shape[0] - 1) / e->shape[0];
        t.nbj_scal = (ny + e->shape[2] - 1) / e->shape[2];
        t.mom0 = mom;
        mom += t.nbj_mom * ((nx + e->shape[1] - 1) / e->shape[1]);
        t.scal0 = scal;
        scal += t.nbj_scal * ((nx + e->shape[3] - 1) / e->shape[3]);
        t.post0 = post;
        post += ((t.g.NY + 3) / 4) * (t.g.NX - 6);
        t.px0 = px;
        if (e->batched_pressure) px += t.g.NY * ((t.nk2 + t.kcx - 1) / t.kcx);
        t.py0 = py;
        if (e->batched_pressure) py += t.ni / 2 + 1;
        t.cp0 = cp;
        cp += 4 * (int)(((long long)t.g.NX * t.g.NYNZ + 1023) / 1024);
    }
    tmom = mom;
    tscal = scal;
    tpost = post;
    tpx = px;
    tpy = py;
    tcp = cp;
}

// offsets of the whole ensemble (the autotuner's launches) and of every group; uploads the tables
// (never during a capture)
int layout(Ensemble *e) {
    offsets(e, e->h, e->tot_mom, e->tot_scal, e->tot_post, e->tot_px, e->tot_py, e->tot_cp);
    DYN_HIP(hipMemcpy(e->d, e->h.data(), sizeof(EnsMember) * e->nmem, hipMemcpyHostToDevice));
    for (EnsGroup &gr : e->groups) {
        gr.h.clear();
        for (int m : gr.idx) gr.h.push_back(e->h[m]);
        offsets(e, gr.h, gr.tot_mom, gr.tot_scal, gr.tot_post, gr.tot_px, gr.tot_py, gr.tot_cp);
        DYN_HIP(hipMemcpy(gr.d, gr.h.data(), sizeof(EnsMember) * gr.h.size(), hipMemcpyHostToDevice));
    }
    return DYN_OK;
}

void free_groups(Ensemble *e) {
    for (size_t g = 0; g < e->groups.size(); g++) {
        EnsGroup &gr = e->groups[g];
        if (gr.d) (void)hipFree(gr.d);
        if (gr.done) (void)hipEventDestroy(gr.done);
        if (g > 0 && gr.stream) (void)hipStreamDestroy(gr.stream);
    }
    e->groups.clear();
}

// split the members into n groups of about equal point counts (largest first, each to the lightest group)
int make_groups(Ensemble *e, int n) {
    free_groups(e);
    n = std::max(1, std::min(n, e->nmem));
    std::vector<int> order(e->nmem);
    for (int m = 0; m < e->nmem; m++) order[m] = m;
    auto pts = [&](int m) { return (long long)e->mem[m]->npts; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return pts(a) > pts(b); });
    e->groups.resize(n);
    std::vector<long long> load(n, 0);
    for (int m : order) {
        const int g = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        e->groups[g].idx.push_back(m);
        load[g] += pts(m);
    }
    for (int g = 0; g < n; g++) {
        EnsGroup &gr = e->groups[g];
        std::sort(gr.idx.begin(), gr.idx.end());
        gr.stream = e->c0()->stream;
        if (g > 0 && hipStreamCreateWithFlags(&gr.stream, hipStreamNonBlocking) != hipSuccess) {
            gr.stream = nullptr;
            set_error("dyn_ensemble: stream creation failed");
            return DYN_E_HIP;
        }
        if (hipEventCreateWithFlags(&gr.done, hipEventDisableTiming) != hipSuccess ||
            hipMalloc(&gr.d, sizeof(EnsMember) * gr.idx.size()) != hipSuccess) {
            set_error("dyn_ensemble: group allocation failed");
            return DYN_E_HIP;
        }
    }
    return layout(e);
}

void drop_graph(Ensemble *e) {
    if (e->exec) (void)hipGraphExecDestroy(e->exec);
    if (e->graph) (void)hipGraphDestroy(e->graph);
    e->exec = nullptr;
    e->graph = nullptr;
}

// swap the current and forecast roles of u, v, w, theta in the ensemble and in every member context
void advance_roles(Ensemble *e) {
    const int fs[4] = {DYN_U, DYN_V, DYN_W, DYN_TH};
    for (int f : fs) {
        std::swap(e->perm[f][DYN_ROLE_CUR], e->perm[f][DYN_ROLE_P]);
        for (Ctx *c : e->mem) std::swap(c->prog[f][DYN_ROLE_CUR], c->prog[f][DYN_ROLE_P]);
    }
}

// runs f(c) with the context's stream replaced by `s` (the batched launchers take their stream from
// the context they are given)
template <typename F>
int on_stream(Ctx *c, hipStream_t s, F &&f) {
    hipStream_t own = c->stream;
    c->stream = s;
    const int rc = f(c);
    c->stream = own;
    return rc;
}

// the launches of one dry RK3 step of every member: each group's passes on its own stream (roles
// rotate on the host between the stages)
int ens_launch_step(Ensemble *e) {
    Ctx *c0 = e->c0();
    const double steps[] = {1. / 3., 0.5, 1.0};
    int rc;
    for (int s = 0; s < 3; s++) {
        const EnsRoles ro = roles_of(e);
        for (EnsGroup &gr : e->groups) {
            const int nm = (int)gr.idx.size();
            rc = on_stream(c0, gr.stream, [&](Ctx *c) {
                return launch_ens_stage(c, gr.d, nm, ro, steps[s], e->shape, gr.tot_mom, gr.tot_scal);
            });
            if (rc) return rc;
            if (e->batched_pressure) {
                rc = on_stream(c0, gr.stream, [&](Ctx *c) {
                    return launch_pressure_ens(c, gr.d, nm, ro, steps[s], gr.tot_px, gr.tot_py, e->lds_x, e->lds_y);
                });
            } else {
                // FP64 members: each member's own solve, on the group's stream
                for (int m : gr.idx) {
                    rc = on_stream(e->mem[m], gr.stream, [&](Ctx *c) { return launch_solve_pressure(c, steps[s], false); });
                    if (rc) break;
                }
            }
            if (rc) return rc;
            rc = on_stream(c0, gr.stream, [&](Ctx *c) { return launch_ens_post(c, gr.d, nm, ro, steps[s], true, gr.tot_post); });
            if (rc) return rc;
        }
        if (s < 2) advance_roles(e);
    }
    const EnsRoles ro = roles_of(e);
    for (EnsGroup &gr : e->groups) {
        const int nm = (int)gr.idx.size();
        rc = on_stream(c0, gr.stream, [&](Ctx *c) { return launch_ens_post(c, gr.d, nm, ro, 0.0, false, gr.tot_post); });
        if (rc) return rc;
        for (int m : gr.idx) {
            Ctx *cm = e->mem[m];
            if (cm->g.rayleigh_level >= cm->g.nz) continue;
            rc = on_stream(cm, gr.stream, [&](Ctx *c) { return launch_rayleigh(c, 1, c->g.nx - 1, 1, c->g.ny - 1, c->g.rayleigh_level); });
            if (rc) return rc;
        }
        rc = on_stream(c0, gr.stream, [&](Ctx *c) { return launch_ens_copy(c, gr.d, nm, ro, gr.tot_cp); });
        if (rc) return rc;
    }
    return DYN_OK;
}

// one dry RK3 step of every member, in the fused form (fused_step of the single context); each group's
// passes go to its own stream, forked from and joined back into the first member's stream (the joins
// follow even when a launch fails, so no work is left on a group's stream)
int ens_step(Ensemble *e) {
    Ctx *c0 = e->c0();
    const int ng = (int)e->groups.size();
    if (ng > 1) {
        DYN_HIP(hipEventRecord(e->fork, c0->stream));
        for (int g = 1; g < ng; g++) DYN_HIP(hipStreamWaitEvent(e->groups[g].stream, e->fork, 0));
    }
    const int rc = ens_launch_step(e);
    for (int g = 1; g < ng; g++) {
        DYN_HIP(hipEventRecord(e->groups[g].done, e->groups[g].stream));
        DYN_HIP(hipStreamWaitEvent(c0->stream, e->groups[g].done, 0));
    }
    if (rc) return rc;
    advance_roles(e);
    return DYN_OK;
}

// time one eager stage-advection launch pair (events on the ensemble stream), in ms
int time_stage(Ensemble *e, float &ms) {
    Ctx *c0 = e->c0();
    hipEvent_t a, b;
    DYN_HIP(hipEventCreate(&a));
    DYN_HIP(hipEventCreate(&b));
    int rc = launch_ens_stage(c0, e->d, e->nmem, roles_of(e), 1.0 / 3.0, e->shape, e->tot_mom, e->tot_scal);  // warm
    DYN_HIP(hipEventRecord(a, c0->stream));
    for (int r = 0; r < 3 && !rc; r++) rc = launch_ens_stage(c0, e->d, e->nmem, roles_of(e), 1.0 / 3.0, e->shape, e->tot_mom, e->tot_scal);
    DYN_HIP(hipEventRecord(b, c0->stream));
    DYN_HIP(hipEventSynchronize(b));
    DYN_HIP(hipEventElapsedTime(&ms, a, b));
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return rc;
}

}  // namespace
}  // namespace dyn

using namespace dyn;

extern "C" {

int dyn_ensemble_create(dyn_ctx *const *members, int n, dyn_ensemble **out) {
    if (!members || n <= 0 || !out) {
        set_error("dyn_ensemble_create: need at least one member");
        return DYN_E_ARG;
    }
    *out = nullptr;
    Ctx *c0 = reinterpret_cast<Ctx *>(members[0]);
    for (int m = 0; m < n; m++) {
        Ctx *c = reinterpret_cast<Ctx *>(members[m]);
        if (!c) {
            set_error("dyn_ensemble_create: null member");
            return DYN_E_ARG;
        }
        if (c->prec != c0->prec || c->g.nz != c0->g.nz) {
            set_error("dyn_ensemble_create: members must share the precision and NZ");
            return DYN_E_ARG;
        }
        if (c->g.microphysics != 0) {
            set_error("dyn_ensemble_create: ensemble members must be dry (microphysics 0)");
            return DYN_E_UNSUPPORTED;
        }
        if (c->turb.on) {
            set_error("dyn_ensemble_create: ensemble members step without turbulence");
            return DYN_E_UNSUPPORTED;
        }
        if (!fused_supported(c) || fused_waves(c) != 1) {
            set_error("dyn_ensemble_create: members need NZ <= 64");
            return DYN_E_UNSUPPORTED;
        }
        for (int k = 0; k < m; k++)
            if (members[k] == members[m]) {
                set_error("dyn_ensemble_create: a context appears twice");
                return DYN_E_ARG;
            }
    }
    Ensemble *e = new Ensemble();
    e->nmem = n;
    e->h.resize(n);
    bool all_batched = c0->prec == DYN_FP32;
    for (int m = 0; m < n; m++) {
        Ctx *c = reinterpret_cast<Ctx *>(members[m]);
        e->mem.push_back(c);
        EnsMember &t = e->h[m];
        memset(&t, 0, sizeof(t));
        t.g = c->geo;
        t.c = c->cols;
        for (int f = 0; f < 7; f++)
            for (int r = 0; r < 3; r++) t.buf[f][r] = c->prog[f][r];
        for (int f = 0; f < 6; f++) t.bar[f] = c->bar[f];
        t.pi = c->pi;
        t.vt = c->vt;
        t.sponge_n = c->sponge_n;
        t.sponge_s = c->sponge_s;
        t.js = c->g.sponge_south;
        t.jn = c->g.sponge_north;
        if (c->g.rayleigh_level < c->g.nz) {
            int rc = rayleigh_table(c, c->g.rayleigh_level);
            if (rc) {
                delete e;
                return rc;
            }
        }
        size_t lx = 0, ly = 0;
        if (all_batched && pressure_ens_fill(c, t, lx, ly)) {
            e->lds_x = std::max(e->lds_x, lx);
            e->lds_y = std::max(e->lds_y, ly);
        } else {
            all_batched = false;
        }
    }
    e->batched_pressure = all_batched;
    for (int f = 0; f < 7; f++)
        for (int r = 0; r < 3; r++) e->perm[f][r] = (unsigned char)r;
    if (c0->prec != DYN_FP32) {
        e->shape[0] = e->shape[2] = 4;
    }
    int rc = ens_prepare(c0->prec == DYN_FP32);
    if (!rc && hipMalloc(&e->d, sizeof(EnsMember) * n) != hipSuccess) {
        set_error("dyn_ensemble_create: table allocation failed");
        rc = DYN_E_HIP;
    }
    if (!rc && hipEventCreateWithFlags(&e->fork, hipEventDisableTiming) != hipSuccess) {
        set_error("dyn_ensemble_create: event creation failed");
        rc = DYN_E_HIP;
    }
    if (!rc) rc = make_groups(e, std::min(kDefaultGroups, n));
    if (rc) {
        free_groups(e);
        if (e->fork) (void)hipEventDestroy(e->fork);
        if (e->d) (void)hipFree(e->d);
        delete e;
        return rc;
    }
    *out = reinterpret_cast<dyn_ensemble *>(e);
    return DYN_OK;
}

int dyn_ensemble_destroy(dyn_ensemble *ens) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_OK;
    (void)hipStreamSynchronize(e->c0()->stream);
    drop_graph(e);
    free_groups(e);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->d) (void)hipFree(e->d);
    delete e;
    return DYN_OK;
}

int dyn_ensemble_set_groups(dyn_ensemble *ens, int ngroups) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e || ngroups < 1) {
        set_error("dyn_ensemble_set_groups: need an ensemble and at least one group");
        return DYN_E_ARG;
    }
    DYN_HIP(hipStreamSynchronize(e->c0()->stream));
    drop_graph(e);
    return make_groups(e, ngroups);
}

int dyn_ensemble_groups(dyn_ensemble *ens) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_E_ARG;
    return (int)e->groups.size();
}

int dyn_ensemble_synchronize(dyn_ensemble *ens) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_E_ARG;
    DYN_HIP(hipStreamSynchronize(e->c0()->stream));
    return DYN_OK;
}

int dyn_ensemble_step_eager(dyn_ensemble *ens, int nsteps) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_E_ARG;
    for (int n = 0; n < nsteps; n++) {
        int rc = ens_step(e);
        if (rc) return rc;
    }
    return DYN_OK;
}

int dyn_ensemble_step(dyn_ensemble *ens, int nsteps) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_E_ARG;
    if (nsteps <= 0) return DYN_OK;
    Ctx *c0 = e->c0();
    const int pairs = nsteps / 2;
    if (pairs > 0) {
        if (!e->exec) {
            unsigned char saved[7][3];
            memcpy(saved, e->perm, sizeof(saved));
            std::vector<std::vector<void *>> saved_prog;
            for (Ctx *c : e->mem) saved_prog.emplace_back(&c->prog[0][0], &c->prog[0][0] + 21);
            DYN_HIP(hipStreamBeginCapture(c0->stream, hipStreamCaptureModeThreadLocal));
            for (Ctx *c : e->mem) c->capturing = true;
            int rc = ens_step(e);
            if (!rc) rc = ens_step(e);
            for (Ctx *c : e->mem) c->capturing = false;
            hipGraph_t graph = nullptr;
            hipError_t err = hipStreamEndCapture(c0->stream, &graph);
            if (rc || err != hipSuccess || memcmp(saved, e->perm, sizeof(saved)) != 0) {
                if (graph) (void)hipGraphDestroy(graph);
                memcpy(e->perm, saved, sizeof(saved));
                for (size_t m = 0; m < e->mem.size(); m++) memcpy(&e->mem[m]->prog[0][0], saved_prog[m].data(), sizeof(void *) * 21);
                if (!rc) set_error(err != hipSuccess ? std::string("hipStreamEndCapture: ") + hipGetErrorString(err)
                                                     : std::string("internal: two ensemble steps did not restore the roles"));
                return rc ? rc : DYN_E_HIP;
            }
            DYN_HIP(hipGraphInstantiate(&e->exec, graph, nullptr, nullptr, 0));
            e->graph = graph;
        }
        for (int p = 0; p < pairs; p++) DYN_HIP(hipGraphLaunch(e->exec, c0->stream));
    }
    if (nsteps % 2) return ens_step(e);
    return DYN_OK;
}

// Launch-shape autotuner: times the stage-advection pair on the members' real grids for every
// (rows per block, march length) of the momentum and the scalar kernel, one kernel at a time, and keeps
// the fastest; the advection kernels write only the forecast buffers, so the state is left as it was.
int dyn_ensemble_autotune(dyn_ensemble *ens, int *shape_out) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e) return DYN_E_ARG;
    drop_graph(e);
    const int tjs[3] = {2, 4, 8}, tis[4] = {8, 16, 32, 64};
    for (int kern = 0; kern < 2; kern++) {
        int best_tj = e->shape[2 * kern], best_ti = e->shape[2 * kern + 1];
        float best = 1e30f;
        for (int tj : tjs)
            for (int ti : tis) {
                e->shape[2 * kern] = tj;
                e->shape[2 * kern + 1] = ti;
                int rc = layout(e);
                float ms = 0;
                if (!rc) rc = time_stage(e, ms);
                if (rc) return rc;
                if (ms < best) {
                    best = ms;
                    best_tj = tj;
                    best_ti = ti;
                }
            }
        e->shape[2 * kern] = best_tj;
        e->shape[2 * kern + 1] = best_ti;
    }
    int rc = layout(e);
    if (rc) return rc;
    if (shape_out) memcpy(shape_out, e->shape, sizeof(e->shape));
    return DYN_OK;
}

int dyn_ensemble_set_shape(dyn_ensemble *ens, const int *shape) {
    Ensemble *e = reinterpret_cast<Ensemble *>(ens);
    if (!e || !shape) return DYN_E_ARG;
    for (int k = 0; k < 4; k++) {
        const bool tj = (k % 2) == 0;
        if (tj ? (shape[k] != 2 && shape[k] != 4 && shape[k] != 8) : (shape[k] < 2 || shape[k] > 512)) {
            set_error("dyn_ensemble_set_shape: rows per block must be 2, 4 or 8 and march lengths 2..512");
            return DYN_E_ARG;
        }
    }
    (void)hipStreamSynchronize(e->c0()->stream);
    drop_graph(e);
    memcpy(e->shape, shape, sizeof(e->shape));
    return layout(e);
}

}  // extern "C"
