"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU).

- The oracle with oracle/sanitize_main.c, which drives every entry point on dry, moist and odd-sized
  states (oracle/Makefile, target sanitize).
- The drop-in layer (model_amd/csrc/refapi/refapi.cpp) inside the reference-shaped executable
  (tests/refshape/Makefile, target asan): loading a state, the reference's initialisation order, and the
  error path a missing device takes (on a CPU-only machine dyn_create reports no HIP device, and the
  executable exits with its drop-in error status, 4).
The sanitizers stop the program at the first invalid access or undefined operation, so a clean exit and
no report pass.  GPU kernels are not sanitized.
"""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")


def _clean(stderr):
    return "AddressSanitizer" not in stderr and "runtime error:" not in stderr


def test_oracle_under_sanitizers():
    exe = os.path.join(ROOT, "oracle", "_build", "oracle_sanitized")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build()")
    r = subprocess.run([exe], capture_output=True, text=True, env=dict(ENV, ASAN_OPTIONS="detect_leaks=1"), timeout=600)
    assert r.returncode == 0 and _clean(r.stderr), r.stdout[-2000:] + r.stderr[-4000:]
    assert "oracle sanitizer run: ok" in r.stdout


def test_drop_in_layer_under_sanitizers(tmp_path):
    exe = os.path.join(ROOT, "tests", "refshape", "_build", "refshape_asan")
    if not os.path.exists(exe):
        pytest.fail(f"{exe} is missing: run __graft_entry__.build()")
    sys.path.insert(0, os.path.dirname(__file__))
    from test_gpu_parity import _write_refshape_input
    from model_amd import configs
    c = configs.supercell(nx=24, ny=24, nz=30)
    _write_refshape_input(c, tmp_path / "in.bin")
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "2"], capture_output=True, text=True,
                       env=ENV, timeout=600)
    assert _clean(r.stderr), r.stderr[-4000:]
    # a machine with a GPU runs the steps (exit 0); without one the drop-in reports it (exit 4)
    assert r.returncode in (0, 4), r.stdout[-2000:] + r.stderr[-2000:]
    if r.returncode == 4:
        assert "drop-in error" in r.stderr
