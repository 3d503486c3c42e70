"""The drop-in link recipe, checked on an executable shaped like the reference's serial build (CPU).

tests/refshape/ holds one stand-in source per reference object (initializer, driver_serial, pressure,
advection, fluxes, kessler, microphysics, boundaries, damping, solver, turbulence), each defining what that object
defines; the entry points the drop-in replaces stop the program if they ever run.  The Makefile
localizes the replaced symbols in those objects with the list libdyncore_refapi's build writes, and
links them against the library (refshape_driver); refshape_control links the same objects without the
localization.  These tests read the symbol tables and the dynamic linker's bindings; the run against
the oracle is in tests/test_gpu_parity.py.
"""
import os
import re
import subprocess

import pytest

from model_amd import build

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
OUT = os.path.join(ROOT, "tests", "refshape", "_build")
DRIVER = os.path.join(OUT, "refshape_driver")
CONTROL = os.path.join(OUT, "refshape_control")
HOST_OBJECTS = ["pressure", "advection", "fluxes", "kessler", "microphysics", "boundaries", "damping", "solver", "turbulence"]
# entry points the reference's time loop calls, one or more per replaced object
MUST_BIND = ["apply_boundary_condition", "apply_boundary_condition_microphysics", "rayleigh_damping", "zero_moisture",
             "run_microphysics", "microphysics_advance", "microphysics_advance_inner", "solve_pressure",
             "initialize_pressure_solver", "integrate_non_hydro", "advect_theta", "advect_microphysics_cell",
             "advance_inner", "advance_outer", "compute_sign_cells", "calculate_diff_tend", "apply_velocity_diffusion",
             "apply_scalar_diffusion", "apply_moisture_diffusion"]
# functions of the replaced objects the library leaves to the reference
KEPT = {"boundaries": ["init_boundaries"], "damping": ["init_damping"], "kessler": ["calculate_rain_fall_speed_kessler"],
        "microphysics": ["init_microphysics", "precip_rate", "calculate_eulerian_fall_speed_rain"],
        "solver": ["switch_array"], "fluxes": ["initialize_flux_cells"], "turbulence": ["init_kmix", "initialize_landsea"]}


def _demangled(path, *flags):
    return subprocess.run(["nm", "-C", *flags, path], capture_output=True, text=True, check=True).stdout


def _replaced():
    names = open(build.REFAPI_REPLACED).read().split()
    dem = subprocess.run(["c++filt"], input="\n".join(names), capture_output=True, text=True, check=True).stdout.split("\n")
    return {d.split("(")[0] for d in dem if d}


def _bindings(exe):
    env = dict(os.environ, LD_BIND_NOW="1", LD_DEBUG="bindings")
    r = subprocess.run([exe, "--load-only"], capture_output=True, text=True, env=env, timeout=60)
    assert r.returncode == 0, r.stderr[-2000:]
    out = {}
    for m in re.finditer(r"binding file (\S+) \[\d+\] to (\S+) \[\d+\]: normal symbol `([^']+)'", r.stderr):
        out.setdefault(m.group(3), []).append((os.path.basename(m.group(1)), os.path.basename(m.group(2))))
    return out


def _mangled_names():
    return {subprocess.run(["c++filt", s], capture_output=True, text=True).stdout.strip().split("(")[0]: s
            for s in open(build.REFAPI_REPLACED).read().split()}


@pytest.fixture(scope="module")
def built():
    for p in (DRIVER, CONTROL, build.REFAPI_REPLACED):
        if not os.path.exists(p):
            pytest.fail(f"{p} is missing: run __graft_entry__.build()")


def test_replaced_list_is_the_header_entry_points(built):
    header = open(os.path.join(ROOT, "include", "dyncore_refapi.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S).split('extern "C"')[0]
    declared = set(re.findall(r"^\s*void\s+([a-z_0-9]+)\s*\(", header, flags=re.M))
    assert _replaced() == declared
    assert set(MUST_BIND) <= declared


@pytest.mark.parametrize("obj", HOST_OBJECTS)
def test_localized_objects_hide_only_the_replaced_definitions(built, obj):
    plain = _demangled(os.path.join(OUT, obj + ".o"), "--defined-only")
    local = _demangled(os.path.join(OUT, obj + ".local.o"), "--defined-only")
    replaced = _replaced()
    for line in plain.splitlines():
        parts = line.split(maxsplit=2)
        if len(parts) < 3 or parts[1] != "T":
            continue
        name = parts[2].split("(")[0]
        sig = re.escape(parts[2])
        if name in replaced:
            assert re.search(rf" t {sig}$", local, flags=re.M), (obj, parts[2])
        else:
            assert re.search(rf" T {sig}$", local, flags=re.M), (obj, parts[2])
    for k in KEPT.get(obj, []):
        assert re.search(rf" T {k}\(", local), (obj, k)


def test_entry_points_bind_to_the_library(built):
    b = _bindings(DRIVER)
    mangled = _mangled_names()
    for name in MUST_BIND:
        got = b.get(mangled[name], [])
        assert got and all(t == "libdyncore_refapi.so" for _, t in got), (name, got)
    # the library's references to the host globals bind to the executable's definitions
    for g in ("us", "ups", "NX", "MICROPHYSICS_OPTION", "jnbuffer", "raydampheight", "landsea", "USE_TURBULENT_STRESS"):
        got = b.get(g, [])
        assert any(src == "libdyncore_refapi.so" and t == "refshape_driver" for src, t in got), (g, got)


def test_without_localization_the_cpu_definitions_win(built):
    """The control link: the stand-in definitions resolve inside the executable, so nothing binds to the
    library -- the failure the recipe's localization step prevents."""
    b = _bindings(CONTROL)
    mangled = _mangled_names()
    for name in MUST_BIND:
        assert not any(t == "libdyncore_refapi.so" for _, t in b.get(mangled[name], [])), name
    undefined = _demangled(CONTROL, "-D", "--undefined-only")
    assert "apply_boundary_condition" not in undefined
    assert "apply_boundary_condition" in _demangled(DRIVER, "-D", "--undefined-only")
