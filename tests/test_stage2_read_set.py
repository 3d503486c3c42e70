"""Which points of stage 2's output buffer the reference's RK3 step reads before it writes them.

The fused step of the HIP library skips the step-end M <- P copy between the steps of one call: M and
the current role then share a buffer, and stage 2 writes into a buffer that holds an older state than
the reference's (whose stage-2 buffer holds the step's starting state, driver_serial.cpp:178-286,
solver.cpp:229-263).  Before stage 2 it copies w whole and the six-column frame of the other
prognostics from the starting state.  That is exact only if stage 2 writes every other point before
anything reads it.  These tests pin that property on the oracle (the reference's algorithm restated in
C): with the rest of the buffer poisoned before stage 2, every field of every role comes out bit for bit
as without the poison; and the smaller sets the copy does not rely on are shown to matter (so the test
fails if the oracle ever stops reading them).
"""
import numpy as np
import pytest

from model_amd import configs, ids
from oracle.pyoracle import Oracle

PROG = list(ids.PROGNOSTIC)
MOIST_ONLY = (ids.QV, ids.QC, ids.QR)


def _run(c, nsteps, poison_region=None, poison=np.nan):
    """The RK3 step as the reference's call sequence; before stage 2, the stage's output buffer of
    every rotated prognostic gets `poison` on the points poison_region(field, NX, NY, NZ) selects."""
    o = Oracle(c)
    NX, NY, NZ = c.nx, c.ny, c.nz
    moist = c.microphysics
    steps = [1 / 3, 0.5, 1.0]
    for _ in range(nsteps):
        for st in range(3):
            if st == 1 and poison_region is not None:
                for f in PROG:
                    if not moist and f in MOIST_ONLY:
                        continue
                    sel = poison_region(f, NX, NY, NZ)
                    if sel is None:
                        continue
                    a = o.get_field(f, ids.ROLE_P)
                    a[sel] = poison
                    o.set_field(f, ids.ROLE_P, a)
            o.advect_uvw(steps[st], 3, NX - 3, 3, NY - 3)
            o.solve_pressure(steps[st])
            o.pgf(steps[st], 1, NX - 1, 1, NY - 1)
            o.advect_theta(steps[st], 3, NX - 3, 3, NY - 3)
            if moist:
                o.advect_moisture(steps[st], 3, NX - 3, 3, NY - 3)
                o.zero_moisture(1, NX - 1, 1, NY - 1)
            o.apply_bc()
            if moist:
                o.apply_bc_micro()
                if st < 2:
                    o.micro_advance_inner()
            if st < 2:
                o.advance_inner()
        if moist:
            o.kessler(1, NX - 1, 1, NY - 1)
            o.apply_bc_micro()
            o.micro_advance_outer()
        o.apply_bc()
        o.rayleigh(1, NX - 1, 1, NY - 1, c.rayleigh_level)
        o.advance_outer()
    return o


def _state(o, c):
    out = []
    for f in PROG + [ids.PI, ids.VT]:
        roles = (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P) if f in PROG else (0,)
        out += [o.get_field(f, r) for r in roles]
    return out


def _same(x, y):
    return all(np.array_equal(a, b, equal_nan=True) for a, b in zip(x, y))


def _outside_copy(f, NX, NY, NZ):
    """Everything the library does not copy before stage 2: fields other than w, inside the frame."""
    if f == ids.W:
        return None
    return np.s_[6:NX - 6, 6:NY - 6, :]


CASES = {
    "moist-random": lambda: configs.small_case(nx=24, ny=20, nz=14, microphysics=1, amplitude=0.5),
    "dry-random-rayleigh": lambda: configs.small_case(nx=16, ny=16, nz=12, microphysics=0, rayleigh_level=7),
    "supercell": lambda: configs.supercell(nx=32, ny=24, nz=24),
    "vortex": lambda: configs.vortex_in_shear(nx=24, ny=24, nz=20),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("poison", [np.nan, 1e30, 0.0], ids=["nan", "huge", "zero"])
def test_stage2_reads_only_the_copied_points_before_writing(case, poison):
    c = CASES[case]()
    ref = _state(_run(c, 3), c)
    got = _state(_run(c, 3, _outside_copy, poison), c)
    assert _same(ref, got)


def test_the_copied_points_are_needed():
    """The two parts of the copy are read: poisoning w's level 1, or the frame of theta, changes the run."""
    c = CASES["supercell"]()
    ref = _state(_run(c, 2), c)
    w_level1 = _state(_run(c, 2, lambda f, NX, NY, NZ: np.s_[6:NX - 6, 6:NY - 6, 1] if f == ids.W else None, 1e30), c)
    assert not _same(ref, w_level1)
    th_frame = _state(_run(c, 2, lambda f, NX, NY, NZ: np.s_[0:6, :, :] if f == ids.TH else None, 1e30), c)
    assert not _same(ref, th_frame)
