"""Closed-form check of the composed RK3 step (driver_serial.cpp:178-286, solver.cpp:229-263).

A single Fourier mode v'(x) = cos(theta i) is advected by a uniform basic-state wind ubar with
u' = w' = theta' = 0 and f = 0.  In rows far from the north/south sponge and ghost rows the only
tendency left is the x flux divergence of v' by ubar, so the step acts on the mode's complex amplitude
as the Wicker-Skamarock RK3 stability polynomial

    G(z) = 1 + z + z^2/2 + z^3/6,   z = -(ubar dt/dx) (1 - e^{-i theta}) S(theta),

with S the symbol of the fifth-order upwind face value of Wicker & Skamarock (2002, MWR 130:2088),
weights (2, -13, 47, 27, -3)/60 on phi_{i-2} .. phi_{i+2} for a positive advecting wind.  Those
weights are the published ones, typed here, not the oracle's interpolant.  The pressure solve sees
divergence only in the boundary rows; its kx = m response decays geometrically in j, and its kx = 0
response (v'^2 carries a mean) is removed by projecting each row onto e^{i theta i}, and a small
amplitude keeps the quadratic terms that reach the interior in later steps below the tolerance.

Run on the oracle (CPU) and on the HIP library (-m gpu).  A step whose stages start from the current
buffer instead of M has amplification (1 + z/3)(1 + z/2)(1 + z) instead, and the test shows it fails.
"""
import numpy as np
import pytest

from helpers import BACKENDS, make_backend, zero_case
from model_amd import ids

NX_INT, NY, NZ = 64, 128, 8
UBAR, DT, DX = 10.0, 40.0, 1000.0
AMP = 1e-3  # small enough that the quadratic terms (v'^2 in the boundary rows) stay below 1e-13
ROWS = slice(NY // 2 - 8, NY // 2 + 8)
WS5 = np.array([2.0, -13.0, 47.0, 27.0, -3.0]) / 60.0  # Wicker & Skamarock (2002), ubar > 0


def _z(theta):
    S = sum(WS5[n] * np.exp(1j * (n - 2) * theta) for n in range(5))
    return -(UBAR * DT / DX) * (1.0 - np.exp(-1j * theta)) * S


def _case(m):
    c = zero_case(nx=NX_INT, ny=NY, nz=NZ, dt=DT, dx=DX, f0=0.0)
    c.fields["ubar"][:] = UBAR
    theta = 2.0 * np.pi * m / NX_INT
    i = np.arange(c.nx) - 3  # interior column i=3 is x index 0; the halos carry the periodic copies
    c.fields["v"] = AMP * np.cos(theta * i)[:, None, None] * np.ones((1, c.ny, c.nz))
    return c, theta


def _amplitude(v, theta):
    """Complex amplitude of e^{i theta i} in every (j, k) column of the interior, rows ROWS."""
    ii = np.arange(NX_INT)
    a = (2.0 / (NX_INT * AMP)) * np.einsum("ijk,i->jk", v[3:3 + NX_INT], np.exp(-1j * theta * ii))
    return a[ROWS, 1:NZ - 1]


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("m", [8, 16, 21])
@pytest.mark.parametrize("nsteps", [1, 3])
def test_rk3_amplification_of_a_fourier_mode(backend, m, nsteps):
    case, theta = _case(m)
    b = make_backend(backend, case)
    b.step(nsteps)
    if backend == "hip":
        b.sync()
    got = _amplitude(b.get_field(ids.V, ids.ROLE_CUR), theta)
    z = _z(theta)
    G = 1 + z + z * z / 2 + z ** 3 / 6
    err = np.abs(got - G ** nsteps).max()
    print(f"{backend} m={m} steps={nsteps}: |G|={abs(G):.6f} arg={np.angle(G):+.6f}, max |A - G^n| = {err:.2e}")
    assert err <= 1e-12


def test_stage_base_on_current_buffer_fails_the_closed_form():
    """The same operators composed with each stage based on the current buffer (M <- current before
    every stage) give (1 + z/3)(1 + z/2)(1 + z), which the closed form above rejects."""
    from oracle.pyoracle import Oracle
    case, theta = _case(8)
    o = Oracle(case)
    NX, NYa = case.nx, case.ny
    for st, step in enumerate((1.0 / 3.0, 0.5, 1.0)):
        for f in (ids.U, ids.V, ids.W, ids.TH):
            o.set_field(f, ids.ROLE_M, o.get_field(f, ids.ROLE_CUR))
        o.advect_uvw(step, 3, NX - 3, 3, NYa - 3)
        o.solve_pressure(step)
        o.pgf(step, 1, NX - 1, 1, NYa - 1)
        o.advect_theta(step, 3, NX - 3, 3, NYa - 3)
        o.apply_bc()
        if st < 2:
            o.advance_inner()
    o.apply_bc()
    o.advance_outer()
    got = _amplitude(o.get_field(ids.V, ids.ROLE_CUR), theta)
    z = _z(theta)
    G = 1 + z + z * z / 2 + z ** 3 / 6
    wrong = (1 + z / 3) * (1 + z / 2) * (1 + z)
    assert np.abs(got - wrong).max() <= 1e-12
    assert np.abs(got - G).max() > 1e-3
