"""The C-ABI libraries load and export every symbol their headers declare (no GPU calls)."""
import ctypes
import os
import re
import subprocess

import pytest

from model_amd import core

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_c_functions(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dyn_[a-z0-9_]+)\s*\(", text)))


def test_capi_library_exports_every_declared_function():
    lib = core.load()
    names = _declared_c_functions("dyncore.h")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n
    assert set(core.EXPORTS) <= set(names)


def test_capi_reports_errors_without_a_device():
    lib = core.load()
    assert lib.dyn_version() == 1
    # argument validation happens before any HIP call
    g = core.DynGrid(70, 64, 40, 1000.0, 250.0, 2.0, 0, 0, 40, 5, 5, 0)  # periodic_ew = 0
    h = ctypes.c_void_p()
    rc = lib.dyn_create(ctypes.byref(g), ctypes.byref(h))
    assert rc == -2 and b"periodic" in lib.dyn_last_error()
    g = core.DynGrid(70, 64, 40, 1000.0, 250.0, 2.0, 0, 1, 39, 5, 5, 0)  # rayleigh level NZ-1
    assert lib.dyn_create(ctypes.byref(g), ctypes.byref(h)) == -1
    assert lib.dyn_advect_uvw(None, 0.5, 3, 10, 3, 10) == -1


def test_refapi_exports_the_reference_entry_points():
    path = core.REFAPI_PATH
    assert os.path.exists(path)
    out = subprocess.run(["nm", "-DC", "--defined-only", path], capture_output=True, text=True, check=True).stdout
    header = open(os.path.join(ROOT, "include", "dyncore_refapi.h")).read()
    header = re.sub(r"/\*.*?\*/", "", header, flags=re.S)
    cxx = header.split('extern "C"')[0]
    decls = re.findall(r"^\s*void\s+([a-z_0-9]+)\s*\(", cxx, flags=re.M)
    assert "solve_pressure" in decls and "advect_uvw_velocity" in decls and "rayleigh_damping" in decls
    for d in set(decls):
        assert re.search(rf" T {d}\(", out), d
    for d in _declared_c_functions("dyncore_refapi.h"):
        assert re.search(rf" T {d}$", out, flags=re.M), d
    # the library defines no host globals and no time loop: the reference's own objects provide them
    for g in ("us", "NX", "MICROPHYSICS_OPTION", "jnbuffer", "raydampheight"):
        assert not re.search(rf" [BDC] {g}$", out, flags=re.M), g
    assert "s_integrate_rk3" not in out


def test_oracle_is_not_reachable_from_the_product():
    # the product package never imports, links or loads the oracle
    for dirpath, _, files in os.walk(os.path.join(ROOT, "model_amd")):
        for fn in files:
            if fn.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, fn)).read()
                assert "oracle" not in text.lower() or fn == "__init__.py" and "oracle" not in text, os.path.join(dirpath, fn)
    ldd = subprocess.run(["ldd", core.LIB_PATH], capture_output=True, text=True).stdout
    assert "oracle" not in ldd


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N above 1 exits with a message (there is no multi-GPU path)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    r = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
