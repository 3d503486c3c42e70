"""The turbulence parameterization (turbulence.cpp, USE_TURBULENT_STRESS) on the device against the
oracle (needs an MI355X).

- Whole turbulent steps (dyn_step, which replays the per-operator sequence with the closure's calls in
  the reference's places, and dyn_step_reference) in the dry and moist forms, with a land-sea mask of
  both kinds of surface, in every role, to 1e-11 relative; the coefficients and tendencies of the last
  step to the same bound.  The moist run makes cloud, so Kessler's saturation flags feed the moist
  Brunt-Vaisala branch.
- The four entry points called one by one on a spun-up state.
- The reference-shaped executable with USE_TURBULENT_STRESS on: init_kmix and initialize_landsea stay
  the reference's, calculate_diff_tend and the apply_* calls bind to the library.
- The FP32 form against the FP64 oracle, and the refusals: the fused step, the adjoint and ensembles.
"""
import os
import subprocess

import numpy as np
import pytest

from helpers import rel_err
from model_amd import configs, ids
from model_amd.core import DYN_FP32, DynCore, DynError, Ensemble
from oracle.pyoracle import Oracle

pytestmark = pytest.mark.gpu

ROLES = (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P)


def _mask(c):
    ls = np.ones((c.nx, c.ny))
    ls[:, : c.ny // 3] = 0.0  # land in the southern third
    return ls


def _case(name):
    """Random states (every coefficient non-zero; the moist one makes saturated points) and the moist
    supercell (the closure active in the sheared bubble, both drag laws under it)."""
    if name == "supercell":
        return configs.supercell(nx=32, ny=32, nz=30)
    moist = name == "random-moist"
    return configs.small_case(nx=24, ny=24, nz=16, microphysics=int(moist), amplitude=0.2 if moist else 0.3,
                              rayleigh_level=12)


def _pair(c, precision=0, landsea=None):
    o, g = Oracle(c), DynCore(c, precision=precision)
    o.set_turbulence(True, landsea)
    g.set_turbulence(True, landsea)
    return o, g


def _prognostic(c):
    return [ids.U, ids.V, ids.W, ids.TH] + ([ids.QV, ids.QC, ids.QR] if c.microphysics else [])


@pytest.mark.parametrize("name", ["random-dry", "random-moist", "supercell"])
@pytest.mark.parametrize("graph", [False, True], ids=["per-operator", "graph"])
def test_turbulent_steps_against_oracle(name, graph):
    c = _case(name)
    moist = c.microphysics != 0
    steps = 16 if name == "supercell" else 10
    ls = _mask(c)
    o, g = _pair(c, landsea=ls)
    print(f"turbulent {name} {c.nx}x{c.ny}x{c.nz} {steps} steps graph={graph}", flush=True)
    o.step(steps)
    (g.step if graph else g.step_reference)(steps)
    g.sync()
    for f in _prognostic(c):
        for role in ROLES:
            e = rel_err(g.get_field(f, role), o.get_field(f, role))
            assert e <= 1e-11, (ids.FIELD_NAMES[f], role, e)
    assert rel_err(g.get_field(ids.PI, 0), o.get_field(ids.PI, 0)) <= 1e-11
    if moist:
        assert rel_err(g.get_field(ids.ACCRAIN), o.get_field(ids.ACCRAIN)) <= 1e-11
    for which in range(9 if moist else 6):
        ref = o.get_turbulence(which)
        e = rel_err(g.get_turbulence(which), ref)
        assert e <= 1e-11, (which, e)
    K = o.get_turbulence(0)
    assert (K > 0).sum() > 100  # the closure is active
    if moist:
        sat = o.get_turbulence(9)
        np.testing.assert_array_equal(g.get_turbulence(9), sat)
        if name == "random-moist":  # saturated pairs of levels take the moist Brunt-Vaisala branch
            assert (sat[:, :, 1:] * sat[:, :, :-1]).sum() > 0
    g.close()


def test_turbulence_changes_the_result():
    """The same run without the closure differs: the comparison above is not of two untouched states."""
    c = _case("random-dry")
    g0, g1 = DynCore(c), DynCore(c)
    g1.set_turbulence(True)
    g0.step(10)
    g1.step(10)
    g0.sync()
    g1.sync()
    assert rel_err(g1.get_field(ids.TH), g0.get_field(ids.TH)) > 1e-6
    g0.close()
    g1.close()


def test_entry_points_one_by_one():
    c = _case("random-moist")
    ls = _mask(c)
    o, g = _pair(c, landsea=ls)
    o.step(6)
    g.step(6)
    I1, J1 = (1, c.nx - 1), (1, c.ny - 1)
    o.calculate_diff_tend(*I1, *J1)
    g.calculate_diff_tend(*I1, *J1)
    for which in range(9):
        assert rel_err(g.get_turbulence(which), o.get_turbulence(which)) <= 1e-11, which
    for name, box in (("apply_velocity_diffusion", (3, c.nx - 3, 3, c.ny - 3)), ("apply_scalar_diffusion", I1 + J1),
                      ("apply_moisture_diffusion", I1 + J1)):
        getattr(o, name)(0.5, *box)
        getattr(g, name)(0.5, *box)
    g.sync()
    for f in _prognostic(c):
        e = rel_err(g.get_field(f, ids.ROLE_P), o.get_field(f, ids.ROLE_P))
        assert e <= 1e-11, (ids.FIELD_NAMES[f], e)
    g.close()


def test_reference_shaped_executable_with_turbulence(tmp_path):
    from test_gpu_parity import _read_refshape_output, _write_refshape_input
    exe = os.path.join(os.path.dirname(__file__), "refshape", "_build", "refshape_driver")
    c = configs.supercell(nx=24, ny=24, nz=30)
    ls = _mask(c)
    steps = 10
    _write_refshape_input(c, tmp_path / "in.bin", turbulence=True, landsea=ls)
    print(f"refshape_driver turbulent supercell {c.nx}x{c.ny}x{c.nz} {steps} steps", flush=True)
    r = subprocess.run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), str(steps)], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = _read_refshape_output(c, tmp_path / "out.bin")
    o = Oracle(c)
    o.set_turbulence(True, ls)
    o.step(steps)
    for f in _prognostic(c):
        for role in ROLES:
            ref = o.get_field(f, role)
            e = rel_err(got[(f, role)].reshape(ref.shape), ref)
            assert e <= 1e-11, (ids.FIELD_NAMES[f], role, e)


def test_fp32_turbulent_steps_against_fp64_oracle():
    """FP32 storage with the closure evaluated in FP32 (the Kessler thermodynamics and the vertical
    solve stay FP64), 16 steps of the moist supercell: the FP32 error of theta, u and w is below a tenth
    of what the closure itself changes, measured as the oracle's difference with and without it."""
    c = _case("supercell")
    ls = _mask(c)
    o, g = _pair(c, precision=DYN_FP32, landsea=ls)
    plain = Oracle(c)
    o.step(16)
    plain.step(16)
    g.step(16)
    g.sync()
    for f in (ids.TH, ids.U, ids.W):
        ref = o.get_field(f)
        err = np.abs(g.get_field(f) - ref).max()
        effect = np.abs(ref - plain.get_field(f)).max()
        print(f"fp32 turbulent supercell 16 steps {ids.FIELD_NAMES[f]}: error {err:.2e}, closure effect {effect:.2e}",
              flush=True)
        assert effect > 0 and err <= 0.1 * effect, (ids.FIELD_NAMES[f], err, effect)
    g.close()


def test_refusals_and_switching_off():
    c = _case("random-dry")
    g = DynCore(c)
    g.set_turbulence(True)
    with pytest.raises(DynError) as e:
        g.step_fused(1)
    assert e.value.code == -2 and "turbulence" in str(e.value)
    with pytest.raises(DynError) as e:
        g.adjoint_record(2, 0)
    assert e.value.code == -2
    with pytest.raises(DynError):
        Ensemble([g])
    # off again: the fused step runs and matches a context that never had the closure, bit for bit
    g.set_turbulence(False)
    h = DynCore(c)
    g.step(3)
    h.step(3)
    g.sync()
    h.sync()
    np.testing.assert_array_equal(g.get_field(ids.TH), h.get_field(ids.TH))
    with pytest.raises(DynError):
        g.get_turbulence(0)
    g.close()
    h.close()
