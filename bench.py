#!/usr/bin/env python3
"""Throughput of the anelastic dynamical core on one MI355X.

Workload: the moist Kessler supercell case at 256x256x64 (NX=262 with the periodic halo, NZ with its two
ghost levels), dt = 2 s, synthetic sounding and bubble built in memory, all fields resident in HBM
before the timed region.  A step is one full RK3 time step (s_integrate_rk3: three stages of advection,
pressure solve and boundaries, then Kessler, boundaries and Rayleigh damping), replayed from a hipGraph.

`value` is the FP32 form's gridpoint-steps/s (FP32 fields and stencils, FP64 vertical solve); the FP64
plain form is measured beside it.  Gridpoints per step = 256*256*64 (halo columns not counted).  Each
form is timed with HIP events on the context's stream around --steps steps, 5 times, each repeat
bracketed by a device synchronisation; the median is reported, with every repeat and its host-clock
time beside it.

The roofline object describes the dominant kernel of the FP32 step: its algorithmic HBM bytes per
launch (field passes x element size x gridpoints processed) divided by its mean launch time, measured
with HIP events on the context's own stream over an eager run of the same step sequence with the whole
stage on one stream.  Its `traffic` (measured HBM bytes per launch) comes from the rocprofv3 counter
summary profiles/hbm_traffic_fp32_supercell256.json, which `--write-traffic` builds from two counter
runs of `--profile-only`; it is used only when the summary was measured on the same kernel sources
(a sha256 over model_amd/csrc), else `traffic` is null and `traffic_source` says why.

Two more measurements ride along (skip them with --no-extras):
  ensemble  32 dry FP32 vortex members, horizontal sizes drawn from {96,128,160,192}^2, nz=40, advanced
            by one launch per pass over all members from a hipGraph (launch shape autotuned), beside the
            same members stepped as separate contexts (each its own stream and graph, so they overlap);
  adjoint   the hand-written adjoint of the dry FP64 step on the 192x192x40 vortex-in-shear case, 360
            steps with a snapshot every 20; `adjoint_forward_ratio` is the gradient's cost, the recording
            forward pass plus the reverse sweep, over one forward run.
The CPU baseline is the oracle on the host cores, in two child processes: all cores (OMP_NUM_THREADS)
and one thread.
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


class HipEvents:
    """hipEventRecord / hipEventElapsedTime on an arbitrary HIP stream (the context's own stream, which
    torch.cuda.Event cannot see), through the HIP runtime's C API."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.hip = ctypes.CDLL("libamdhip64.so")
        self.hip.hipEventCreate.argtypes = [ctypes.POINTER(ctypes.c_void_p)]
        self.hip.hipEventRecord.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventSynchronize.argtypes = [ctypes.c_void_p]
        self.hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), ctypes.c_void_p, ctypes.c_void_p]
        self.hip.hipEventDestroy.argtypes = [ctypes.c_void_p]

    def _check(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed with HIP error {rc}")

    def time(self, stream, fn):
        """ms between events recorded on `stream` before and after fn() enqueues its work."""
        a, b = self.ct.c_void_p(), self.ct.c_void_p()
        self._check(self.hip.hipEventCreate(self.ct.byref(a)), "hipEventCreate")
        self._check(self.hip.hipEventCreate(self.ct.byref(b)), "hipEventCreate")
        self._check(self.hip.hipEventRecord(a, stream), "hipEventRecord")
        fn()
        self._check(self.hip.hipEventRecord(b, stream), "hipEventRecord")
        self._check(self.hip.hipEventSynchronize(b), "hipEventSynchronize")
        ms = self.ct.c_float()
        self._check(self.hip.hipEventElapsedTime(self.ct.byref(ms), a, b), "hipEventElapsedTime")
        self.hip.hipEventDestroy(a)
        self.hip.hipEventDestroy(b)
        return float(ms.value)


REPEATS = 5


def time_form(case, precision, steps, warmup, stage_streams):
    """K steps timed REPEATS times, each bracketed by a device synchronisation on both sides, with HIP
    events on the context's stream (the graphs' fork and join bring every branch back to it) and with
    the host clock; returns the context, the median event ms per step and the per-repeat records."""
    import statistics
    from model_amd.core import DYN_TUNE_STAGE_STREAMS, DynCore
    g = DynCore(case, precision=precision)
    g.tune_set(DYN_TUNE_STAGE_STREAMS, stage_streams)
    g.step(warmup)
    g.sync()
    ev = HipEvents()
    event_ms, wall_ms = [], []
    for _ in range(REPEATS):
        g.sync()
        t0 = time.perf_counter()
        event_ms.append(ev.time(g.stream, lambda: g.step(steps)) / steps)
        g.sync()
        wall_ms.append((time.perf_counter() - t0) * 1e3 / steps)
    return g, statistics.median(event_ms), {"event_ms_per_step": [round(x, 4) for x in event_ms],
                                            "wall_ms_per_step": [round(x, 4) for x in wall_ms]}


def kernel_profile(g, steps, fused, shape=None):
    """Per-kernel HIP-event timings of the same step sequence the graph replays, launched eagerly, with
    the whole stage on one stream: the scalar half, which the step runs beside the pressure solve, would
    otherwise share the GPU with the kernels it overlaps and its duration would not be its own.  On one
    stream the stage kernels take the library's one-stream launch shapes; `shape` (a dict), when given,
    receives them."""
    from model_amd.core import DYN_TUNE_STAGE_STREAMS
    before = g.tune_get(DYN_TUNE_STAGE_STREAMS)
    g.tune_set(DYN_TUNE_STAGE_STREAMS, 1)
    if shape is not None:
        shape.update(launch_shape(g))
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


def launch_shape(g):
    """Rows per block and march length of the two stage kernels as the context launches them now."""
    from model_amd.core import DYN_TUNE_MOMENTUM_TI, DYN_TUNE_MOMENTUM_TJ, DYN_TUNE_SCALAR_TI, DYN_TUNE_SCALAR_TJ
    return {"momentum_rows": g.tune_get(DYN_TUNE_MOMENTUM_TJ), "momentum_march": g.tune_get(DYN_TUNE_MOMENTUM_TI),
            "scalar_rows": g.tune_get(DYN_TUNE_SCALAR_TJ), "scalar_march": g.tune_get(DYN_TUNE_SCALAR_TI)}


def roofline(rows, elt, moist):
    name, n, ms, units = max(rows, key=lambda r: r[2])
    passes = PASSES.get(name)
    if isinstance(passes, dict):
        passes = passes[moist]
    per_launch_units = units / n
    avg_ms = ms / n
    bytes_per_launch = passes * elt * per_launch_units
    achieved = bytes_per_launch / (avg_ms * 1e-3) / 1e9
    traffic, traffic_status = measured_traffic(name)
    return {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
            "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": traffic, "traffic_source": traffic_status,
            "avg_launch_ms": round(avg_ms, 5),
            "algorithmic_bytes_per_launch": int(bytes_per_launch), "field_passes_per_point": passes,
            "points_per_launch": int(per_launch_units)}


# timing-record name (dyn_timing_report) -> base name of the kernel the record times
KERNEL_SYMBOL = {
    "adv_scalars": "k_adv_scalars", "adv_momentum": "k_adv_momentum", "p1_div_rfft_x": "k_div_rfft_x", "p2_ypass": "k_ypass",
    "p5_irfft_x_out": "k_irfft_x_out", "post_uvw": "k_post_uvw", "post_scalars": "k_post_scalars_shell", "kessler": "k_kessler",
    "precip": "k_precip", "advance_copy": "k_copy_fields",
}
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "hbm_traffic_fp32_supercell256.json")
CSRC = os.path.join(ROOT, "model_amd", "csrc")


def csrc_digest():
    """sha256 over the kernel sources (every file under model_amd/csrc, sorted by path): a counter
    summary is only used for the build it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for dirpath, _, files in sorted(os.walk(CSRC)):
        for fn in sorted(files):
            if fn.endswith((".hip", ".hpp", ".cpp", ".h")):
                path = os.path.join(dirpath, fn)
                h.update(os.path.relpath(path, CSRC).encode())
                with open(path, "rb") as fh:
                    h.update(fh.read())
    return h.hexdigest()


def kernel_base_name(symbol):
    """k_adv_scalars for 'void dyn::(anonymous namespace)::k_adv_scalars<float, 8, true, 1>(dyn::Geo, ...)'."""
    sym = symbol.replace("(anonymous namespace)::", "")
    if sym.startswith("void "):
        sym = sym[5:]
    cut = min([i for i in (sym.find("<"), sym.find("(")) if i >= 0] or [len(sym)])
    return sym[:cut].split("::")[-1]


def measured_traffic(kernel):
    """(HBM bytes per launch of `kernel`, status) from the committed counter summary: FETCH_SIZE +
    WRITE_SIZE, each scaled by what it reads for a copy kernel of known size.  The summary must have been
    measured on these kernel sources (same csrc digest) and hold exactly one instantiation of the kernel,
    else the bytes are None."""
    if not os.path.exists(TRAFFIC_PROFILE):
        return None, "no counter summary"
    with open(TRAFFIC_PROFILE) as f:
        data = json.load(f)
    if data.get("csrc_sha256") != csrc_digest():
        return None, "stale: the counter summary was measured on other kernel sources"
    base = KERNEL_SYMBOL.get(kernel, kernel)
    hits = [row for name, row in data.get("kernels", {}).items() if kernel_base_name(name) == base]
    if len(hits) != 1:
        return None, f"{len(hits)} instantiations of {base} in the summary"
    return int(hits[0]["bytes"]), "measured on these sources"


def write_traffic_summary(fetch_csv, write_csv, copy_bytes):
    """Build the counter summary bench.py reads from two rocprofv3 --pmc runs (FETCH_SIZE, then
    WRITE_SIZE; counter_collection.csv), averaged per kernel symbol over its dispatches, in KiB.  Each
    counter is scaled by what it reads for k_copy_fields, whose launches in the profiled command are the
    7-field M <- P copy of copy_bytes bytes each way."""
    import csv
    import statistics

    def per_kernel(path, counter):
        vals = {}
        with open(path) as fh:
            for row in csv.DictReader(fh):
                if row.get("Counter_Name") != counter:
                    continue
                vals.setdefault(row["Kernel_Name"], []).append(float(row["Counter_Value"]))
        return {k: statistics.mean(v) for k, v in vals.items()}

    fetch, write = per_kernel(fetch_csv, "FETCH_SIZE"), per_kernel(write_csv, "WRITE_SIZE")
    copy = [k for k in fetch if kernel_base_name(k) == "k_copy_fields"]
    if len(copy) != 1:
        raise SystemExit(f"expected one k_copy_fields symbol, found {copy}")
    fs = copy_bytes / 1024.0 / fetch[copy[0]]
    ws = copy_bytes / 1024.0 / write[copy[0]]
    kernels = {}
    for k in sorted(set(fetch) & set(write)):
        kernels[k] = {"fetch_kib": round(fetch[k], 1), "write_kib": round(write[k], 1),
                      "bytes": int(round((fetch[k] * fs + write[k] * ws) * 1024))}
    out = {"csrc_sha256": csrc_digest(), "copy_kernel": copy[0], "copy_bytes_each_way": copy_bytes,
           "fetch_scale": fs, "write_scale": ws, "kernels": kernels}
    with open(TRAFFIC_PROFILE, "w") as fh:
        json.dump(out, fh, indent=1)
    print(json.dumps(out, indent=1))


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


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(budget_s):
    """The oracle (the C restatement of the reference serial path, FP64) on the C2 case, OpenMP over
    i-slabs with OMP_NUM_THREADS threads: one warm-up step, then as many steps as fit the budget."""
    from model_amd import configs
    from oracle.pyoracle import Oracle
    case = configs.supercell()
    o = Oracle(case, variant="omp")
    t0 = time.perf_counter()
    o.step(1)  # first step: also faults the pages in
    first = time.perf_counter() - t0
    n = max(1, min(20, int((budget_s - first) / max(first, 1e-3))))
    t0 = time.perf_counter()
    o.step(n)
    dt = time.perf_counter() - t0
    return {"value": round(case.interior_points * n / dt, 1), "steps": n, "seconds": round(dt, 2)}


def cpu_baseline():
    """Two legs in child processes (the OpenMP thread count is fixed when a process starts): every host
    core the box gives (OMP_NUM_THREADS), the stand-in for the reference's MPI build, and one thread, the
    reference's serial build."""
    import subprocess
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    legs = {}
    for n, budget in ((threads, 10.0), (1, 14.0)):
        env = dict(os.environ, OMP_NUM_THREADS=str(n))
        r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", str(budget)], env=env,
                           capture_output=True, text=True, timeout=300)
        if r.returncode != 0:
            raise RuntimeError("CPU baseline leg failed: " + r.stderr[-1000:])
        legs[n] = json.loads(r.stdout.strip().splitlines()[-1])
    many, one = legs[threads], legs[1]
    return {"value": many["value"], "unit": "gridpoint-steps/s", "cores": threads, "kind": "port",
            "cpu_model": cpu_model(),
            "sample": f"{many['steps']} steps of the same 256x256x64 moist case after 1 warm-up step, FP64 oracle "
                      f"(C restatement of the reference serial path), OpenMP over i-slabs, {many['seconds']} s",
            "single_thread": {"value": one["value"], "cores": 1,
                              "sample": f"{one['steps']} steps after 1 warm-up step, one thread, {one['seconds']} s"}}


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
    ap.add_argument("--profile-only", action="store_true",
                    help="only the eager per-kernel pass of the main form (the command the counter summaries use)")
    ap.add_argument("--cpu-leg", type=float, default=None, help=argparse.SUPPRESS)
    ap.add_argument("--write-traffic", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="build profiles/hbm_traffic_fp32_supercell256.json from two rocprofv3 --pmc runs of "
                         "--profile-only --form fp32 --profile-steps 1")
    args = ap.parse_args()
    if args.cpu_leg is not None:
        print(json.dumps(cpu_leg(args.cpu_leg)))
        return 0
    if args.write_traffic:
        from model_amd import configs
        c = configs.supercell()
        write_traffic_summary(*args.write_traffic, copy_bytes=7 * c.nx * c.ny * c.nz * 4)
        return 0
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
    from model_amd.core import DYN_TUNE_STAGE_STREAMS, DynCore
    if args.profile_only:
        for name, prec in forms[:1]:
            handles[name] = DynCore(case, precision=prec)
    for name, prec in ([] if args.profile_only else forms):
        g, ms, reps = time_form(case, prec, args.steps, args.warmup, args.stage_streams)
        res[name] = {"ms_per_step": round(ms, 4), "value": round(npts / (ms * 1e-3), 1),
                     "stage_streams": g.tune_get(DYN_TUNE_STAGE_STREAMS), **reps}
        handles[name] = g
        log(f"bench: {name} {ms:.3f} ms/step (median of {REPEATS} x {args.steps} steps), "
            f"{res[name]['value']:.4g} gridpoint-steps/s")
    main_form = forms[0][0]
    g = handles[main_form]
    step_shape, profile_shape = launch_shape(g), {}
    rows = kernel_profile(g, args.profile_steps, fused=True, shape=profile_shape)
    for r in sorted(rows, key=lambda r: -r[2]):
        log(f"bench: kernel {r[0]:16s} launches {r[1]:5d} total {r[2]:9.3f} ms avg {r[2] / r[1] * 1e3:8.1f} us")
    roof = roofline(rows, 4 if main_form == "fp32" else 8, case.microphysics)
    # the kernel durations come from the one-stream sequence and its launch shapes; the timed step runs
    # the scalar half beside the pressure solve with its own
    roof["measured_launch_shape"] = profile_shape
    roof["step_launch_shape"] = step_shape
    if args.profile_only:
        log(f"bench: dominant kernel {roof['kernel']} {roof['achieved']} GB/s, traffic {roof['traffic']}")
        return 0
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
        "timing": f"HIP events on the context's stream around {args.steps} steps, median of {REPEATS} repeats, "
                  f"each bracketed by a device synchronisation",
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
        # the gradient's cost: the recording forward pass plus the reverse sweep, over one forward run
        out["adjoint_forward_ratio"] = out["adjoint"]["record_plus_backward_over_forward"]
        log(f"bench: adjoint (record + backward)/forward {out['adjoint_forward_ratio']}")
    if not args.no_cpu:
        log("bench: timing the CPU restatement (oracle) on a bounded sample")
        out["cpu_baseline"] = cpu_baseline()
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
