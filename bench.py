#!/usr/bin/env python3
"""Throughput of the anelastic dynamical core on one MI355X.

Workload: the moist Kessler supercell case at 256x256x64 (NX=262 with the periodic halo, NZ with its two
ghost levels), dt = 2 s, synthetic sounding and bubble built in memory, all fields resident in HBM
before the timed region.  A step is one full RK3 time step (s_integrate_rk3: three stages of advection,
pressure solve and boundaries, then Kessler, boundaries and Rayleigh damping), replayed from a hipGraph.

`value` is the FP32 form's gridpoint-steps/s (FP32 fields and stencils, FP64 vertical solve); the FP64
plain form is measured beside it.  Gridpoints per step = 256*256*64 (halo columns not counted).

The roofline object describes the dominant kernel of the FP32 step: its algorithmic HBM bytes per
launch (field passes x element size x gridpoints processed) divided by its mean launch time, measured
with HIP events on the context's own stream over an instrumented run of the same step sequence.  Its
`traffic` (measured HBM bytes per launch) comes from the rocprofv3 counter summary committed under
profiles/ when one exists for that kernel.

Two more measurements ride along (skip them with --no-extras):
  ensemble  32 dry FP32 vortex members, horizontal sizes drawn from {96,128,160,192}^2, nz=40, advanced
            by one launch per pass over all members from a hipGraph (launch shape autotuned), beside the
            same members stepped as separate contexts (each its own stream and graph, so they overlap);
  adjoint   the hand-written adjoint of the dry FP64 step on the 192x192x40 vortex-in-shear case, 360
            steps with a snapshot every 20: backward time / forward time.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "gridpoint-steps/s at 256²×64 FP32 vs FP64 on 1 GPU; adjoint/forward cost ratio"
PEAK_HBM_GBS = 8000.0

# algorithmic field passes per gridpoint processed, per kernel (one read of each input field and one write of
# each output field; a half-spectrum complex array counts as one real-equivalent field)
PASSES = {
    "advect_uvw": {0: 13, 1: 16}, "advect_theta": 10, "advect_moisture": 16, "fall_speed": 2, "zero_moisture": 7,
    "pgf": 7, "p1_div_fft_x": 4, "p2_fft_y": 2, "p3_thomas": 2, "p4_ifft_y": 2, "p5_ifft_x_out": 2, "kessler": 12,
    "rayleigh": 8, "advance_copy": 2,
    # fused stage, momentum kernel: U V W UB VB WB cur/basic, UM VM WM, theta' (+ qv qc qr) read; UP VP WP written
    "adv_momentum": {0: 6 + 3 + 1 + 3, 1: 6 + 3 + 4 + 3},
    # scalar kernel: U V W UB VB WB, theta' thetabar (+ qv qbar qc qr), M buffers read; forecasts (+ VT) written
    "adv_scalars": {0: 6 + 2 + 1 + 1, 1: 6 + 6 + 4 + 5},
    "p1_div_rfft_x": 4, "p2_ypass": 2, "p5_irfft_x_out": 2,
    "post_uvw": 7, "post_scalars": 1, "zero_moisture_shell": 1,
}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def time_form(case, precision, steps, warmup, stage_streams):
    from model_amd.core import DYN_TUNE_STAGE_STREAMS, DynCore
    g = DynCore(case, precision=precision)
    g.tune_set(DYN_TUNE_STAGE_STREAMS, stage_streams)
    g.step(warmup)
    g.sync()
    t0 = time.perf_counter()
    g.step(steps)
    g.sync()
    t1 = time.perf_counter()
    return g, (t1 - t0) * 1e3 / steps


def kernel_profile(g, steps, fused):
    """Per-kernel HIP-event timings of the same step sequence the graph replays, launched eagerly, with
    the whole stage on one stream: the scalar half, which the step runs beside the pressure solve, would
    otherwise share the GPU with the kernels it overlaps and its duration would not be its own."""
    from model_amd.core import DYN_TUNE_STAGE_STREAMS
    before = g.tune_get(DYN_TUNE_STAGE_STREAMS)
    g.tune_set(DYN_TUNE_STAGE_STREAMS, 1)
    g.timing(1)
    (g.step_fused if fused else g.step_reference)(steps)
    g.sync()
    g.tune_set(DYN_TUNE_STAGE_STREAMS, before)
    import ctypes
    buf = ctypes.create_string_buffer(1 << 16)
    g.timing_report(buf, len(buf))
    g.timing(0)
    rows = []
    for line in buf.value.decode().strip().splitlines():
        name, n, ms, units = line.split()
        rows.append((name, int(n), float(ms), float(units)))
    return rows


def roofline(rows, elt, moist):
    name, n, ms, units = max(rows, key=lambda r: r[2])
    passes = PASSES.get(name)
    if isinstance(passes, dict):
        passes = passes[moist]
    per_launch_units = units / n
    avg_ms = ms / n
    bytes_per_launch = passes * elt * per_launch_units
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": measured_traffic(name), "avg_launch_ms": round(avg_ms, 5),
            "algorithmic_bytes_per_launch": int(bytes_per_launch), "field_passes_per_point": passes,
            "points_per_launch": int(per_launch_units)}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed counter summary (FETCH_SIZE + WRITE_SIZE,
    scaled by a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_fp32_supercell256.json")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        data = json.load(f)
    for name, row in data.get("kernels", {}).items():
        if name.startswith(KERNEL_SYMBOL.get(kernel, kernel)):
            return int(row["bytes"])
    return None


# timing-record name -> kernel symbol prefix in the rocprofv3 summaries
KERNEL_SYMBOL = {"adv_scalars": "k_adv_scalars", "adv_momentum": "k_adv_momentum", "p2_ypass": "k_ypass",
                 "p1_div_rfft_x": "k_div_rfft_x", "p5_irfft_x_out": "k_irfft_x_out", "post_uvw": "k_post_uvw",
                 "post_scalars": "k_post_scalars"}


def ensemble_bench(steps, warmup):
    import numpy as np
    from model_amd import configs
    from model_amd.core import DYN_FP32, DynCore, Ensemble
    rng = np.random.Generator(np.random.PCG64(20261202 + 3))
    sizes = rng.choice([96, 128, 160, 192], size=(32, 2))
    cases = [configs.vortex_in_shear(nx=int(a), ny=int(b), nz=40, seed_offset=100 + m) for m, (a, b) in enumerate(sizes)]
    members = [DynCore(c, precision=DYN_FP32) for c in cases]
    pts = sum(c.interior_points for c in cases)
    ens = Ensemble(members)
    shape = ens.autotune()
    groups = ens.groups
    ens.step(warmup)
    ens.sync()
    t0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    t_ens = time.perf_counter() - t0
    for m in members:  # each member alone: its own stream and graph, launched back to back (they overlap)
        m.step(warmup)
    for m in members:
        m.sync()
    t0 = time.perf_counter()
    for m in members:
        m.step(steps)
    for m in members:
        m.sync()
    t_seq = time.perf_counter() - t0
    ens.close()
    for m in members:
        m.close()
    return {"members": len(cases), "sizes": [[int(a), int(b)] for a, b in sizes], "nz": 40, "dtype": "f32",
            "gridpoints_per_step": pts, "steps": steps, "ms_per_step": round(t_ens * 1e3 / steps, 4),
            "value": round(pts * steps / t_ens, 1), "unit": "gridpoint-steps/s",
            "members_one_by_one": round(pts * steps / t_seq, 1), "batched_over_one_by_one": round(t_seq / t_ens, 3),
            "shape": {"momentum_tj": shape[0], "momentum_ti": shape[1], "scalar_tj": shape[2], "scalar_ti": shape[3]},
            "stream_groups": groups}


def adjoint_bench(steps=360, every=20):
    from model_amd import configs
    from model_amd.adjoint import backward, record, set_initial, vorticity_weights
    from model_amd.core import DynCore
    c = configs.vortex_in_shear()
    core = DynCore(c)
    x0 = [c.fields["u"], c.fields["v"], c.fields["w"], c.fields["th"]]
    Wu, Wv = vorticity_weights(c)
    core.step(4)
    core.sync()
    set_initial(core, x0)
    t0 = time.perf_counter()
    core.step(steps)
    core.sync()
    t_fwd = time.perf_counter() - t0
    runs = []
    for ev in (every, 0):  # a fixed interval, then the longest one that fits in half of free HBM
        set_initial(core, x0)
        t0 = time.perf_counter()
        used = record(core, steps, ev)
        t_rec = time.perf_counter() - t0
        t0 = time.perf_counter()
        backward(core, [Wu, Wv, None, None])
        t_bwd = time.perf_counter() - t0
        core.L.dyn_adjoint_release(core.h)
        runs.append({"checkpoint_every": used, "record_s": round(t_rec, 4), "backward_s": round(t_bwd, 4),
                     "backward_over_forward": round(t_bwd / t_fwd, 3),
                     "record_plus_backward_over_forward": round((t_rec + t_bwd) / t_fwd, 3)})
    core.close()
    return {"workload": "dry vortex in shear 192x192x40, dx=20 km, dz=500 m, dt=60 s, FP64", "steps": steps,
            "forward_s": round(t_fwd, 4), **runs[0], "hbm_resident": runs[1]}


def cpu_baseline(case, budget_s=20.0):
    """The oracle (the C restatement of the reference serial path) on the host cores, OpenMP over i-slabs."""
    from oracle.pyoracle import Oracle
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    o = Oracle(case, variant="omp")
    t0 = time.perf_counter()
    o.step(1)  # first step: also faults the pages in
    first = time.perf_counter() - t0
    n = max(1, min(20, int((budget_s - first) / max(first, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    rate = case.interior_points * n / dt
    return {"value": round(rate, 1), "unit": "gridpoint-steps/s", "cores": threads, "kind": "port",
            "sample": f"{n} steps of the same 256x256x64 moist case (after 1 warm-up step), FP64, "
                      f"OpenMP over i-slabs, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--profile-steps", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--form", choices=["both", "fp32", "fp64"], default="both")
    ap.add_argument("--no-extras", action="store_true", help="skip the ensemble and adjoint measurements")
    ap.add_argument("--stage-streams", type=int, choices=[0, 1, 2, 3, 4], default=0,
                    help="0: the library's default; 2: a stage's scalar half runs on a second stream beside the "
                         "pressure solve, forked after the momentum kernel; 3: forked after the solve's first "
                         "pass; 4: forked after the momentum kernel, queued after the first pass; 1: the whole "
                         "step on one stream (per-kernel profiles)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the dynamical core here has no multi-GPU path; run with --gpus 1",
              file=sys.stderr)
        return 2

    from model_amd import configs
    from model_amd.core import DYN_FP32, DYN_FP64

    case = configs.supercell()
    npts = case.interior_points
    log(f"bench: workload {case.name} NX={case.nx} NY={case.ny} NZ={case.nz}, {npts} gridpoints per step")
    res = {}
    forms = [("fp32", DYN_FP32), ("fp64", DYN_FP64)] if args.form == "both" else \
        [(args.form, DYN_FP32 if args.form == "fp32" else DYN_FP64)]
    handles = {}
    for name, prec in forms:
        g, ms = time_form(case, prec, args.steps, args.warmup, args.stage_streams)
        from model_amd.core import DYN_TUNE_STAGE_STREAMS
        res[name] = {"ms_per_step": round(ms, 4), "value": round(npts / (ms * 1e-3), 1),
                     "stage_streams": g.tune_get(DYN_TUNE_STAGE_STREAMS)}
        handles[name] = g
        log(f"bench: {name} {ms:.3f} ms/step, {res[name]['value']:.4g} gridpoint-steps/s")
    main_form = forms[0][0]
    g = handles[main_form]
    rows = kernel_profile(g, args.profile_steps, fused=True)
    for r in sorted(rows, key=lambda r: -r[2]):
        log(f"bench: kernel {r[0]:16s} launches {r[1]:5d} total {r[2]:9.3f} ms avg {r[2] / r[1] * 1e3:8.1f} us")
    roof = roofline(rows, 4 if main_form == "fp32" else 8, case.microphysics)
    out = {
        "metric": METRIC,
        "value": res[main_form]["value"],
        "unit": "gridpoint-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": res[main_form]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if main_form == "fp32" else "f64",
        "data": "synthetic",
        "config": {"workload": "moist Kessler supercell, 256x256x64 (NX=262 incl. halo), dx=1 km, dz=250 m, dt=2 s",
                   "nx": 256, "ny": 256, "nz": 64, "form": main_form, "step": "hipGraph replay of s_integrate_rk3",
                   "stage_streams": res[main_form]["stage_streams"]},
        "forms": res,
        "fp32_over_fp64": round(res["fp32"]["value"] / res["fp64"]["value"], 3) if "fp64" in res and "fp32" in res else None,
        "adjoint_forward_ratio": None,
        "roofline": roof,
    }
    if not args.no_extras:
        log("bench: ensemble (32 members)")
        out["ensemble"] = ensemble_bench(max(10, args.steps // 5), 6)
        log(f"bench: ensemble {out['ensemble']['value']:.4g} gridpoint-steps/s, "
            f"{out['ensemble']['batched_over_one_by_one']}x the members one by one")
        log("bench: adjoint (192x192x40, 360 steps)")
        out["adjoint"] = adjoint_bench()
        out["adjoint_forward_ratio"] = out["adjoint"]["backward_over_forward"]
        log(f"bench: adjoint backward/forward {out['adjoint_forward_ratio']}")
    if not args.no_cpu:
        log("bench: timing the CPU restatement (oracle) on a bounded sample")
        out["cpu_baseline"] = cpu_baseline(case)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
