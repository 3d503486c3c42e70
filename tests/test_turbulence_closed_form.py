"""The turbulence parameterization (turbulence.cpp) against closed forms, on the oracle and the HIP library.

- Mixing coefficient: for a uniform vertical shear U' in a stratified column the deformation is S = U'
  and the coefficient is the Smagorinsky-Lilly form with the Richardson-number reduction,
  K = (c_s l)^2 sqrt(S^2 - N^2 / Pr), c_s = 0.25, l = min(dz, 200 m), Pr = 1/3, zero when Ri = N^2/S^2 >=
  Pr, bounded by 0.15 l^2/dt; the horizontal coefficient on the level below is (dx/4)^2 times the mean
  of K/l^2 of the two levels, bounded by 0.05 dx^2/dt (turbulence.cpp:85-147, 737-831).
- Scalar diffusion: with the coefficients those of uniform shear, the theta tendency of a sinusoid in x
  is the conservative second difference 3 K_H (th(i+1) - 2 th(i) + th(i-1)) / dx^2 (scalar diffusivity
  Pr^-1 = 3 times the momentum one), and vertical diffusion conserves the column integral (zero flux at
  the ground and the lid).
- Momentum: in uniform shear with uniform K the interior friction vanishes; at the lowest level it is the
  bulk surface drag -C_D |V| u / dz with the ocean's C_D = 1.1e-3 + 4e-5 |V| (land 3e-3), and at the top
  the stress-free lid gives -K U' / dz.
"""
import numpy as np
import pytest

from helpers import BACKENDS, make_backend, zero_case
from model_amd import ids

GRAV = 9.8
CS, LVMAX, PR_INV = 0.25, 200.0, 3.0


def _shear_case(ushear, n2, nx=16, ny=12, nz=14, u0=0.0, theta0=300.0):
    c = zero_case(nx=nx, ny=ny, nz=nz, dx=1000.0, dz=250.0, dt=2.0)
    z = c.columns["zsu"]
    tb = theta0 * np.exp(n2 * z / GRAV)
    c.columns["tb"] = tb
    c.columns["tbv"] = tb.copy()
    c.fields["u"][:] = (u0 + ushear * z)[None, None, :]
    return c


def _backend(kind, c, landsea=None):
    b = make_backend(kind, c)
    b.set_turbulence(True, landsea)
    b.calculate_diff_tend(1, c.nx - 1, 1, c.ny - 1)
    if kind == "hip":
        b.sync()
    return b


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("ushear,n2", [(4e-3, 1e-6), (4e-3, 1e-5), (1e-2, 1e-5), (2e-2, 0.0), (20.0, 1e-5)])
def test_mixing_coefficient_smagorinsky_lilly(backend, ushear, n2):
    c = _shear_case(ushear, n2)
    b = _backend(backend, c)
    K, KH = b.get_turbulence(0), b.get_turbulence(1)
    z, tb = c.columns["zsu"], c.columns["tb"]
    l2 = min(c.dz, LVMAX) ** 2
    raw = np.zeros(c.nz)
    for k in range(2, c.nz - 1):
        N2 = 2 * GRAV / (tb[k] + tb[k - 1]) * (tb[k] - tb[k - 1]) / (z[k] - z[k - 1])
        S2 = ushear ** 2
        raw[k] = (CS ** 2) * l2 * np.sqrt(S2 - PR_INV * N2) if S2 > PR_INV * N2 else 0.0
    kmax = 0.15 * l2 / c.dt
    expect = np.minimum(raw, kmax)
    for k in range(2, c.nz - 1):
        np.testing.assert_allclose(K[2:-2, 2:-2, k], expect[k], rtol=1e-12, atol=1e-300)
    # the ground and the lid
    assert not K[1:-1, 1:-1, 1].any() and not K[1:-1, 1:-1, c.nz - 1].any()
    # horizontal on level k: (dx/4)^2 times the mean of K/l^2 of levels k (bounded) and k+1 (before its
    # bound), itself bounded by 0.05 dx^2/dt
    for k in range(1, c.nz - 2):
        hk = 0.5 * (expect[k] / l2 + raw[k + 1] / l2) * (c.dx / 4) ** 2 if raw[k + 1] > 0 else 0.5 * expect[k] / l2 * (c.dx / 4) ** 2
        np.testing.assert_allclose(KH[2:-2, 2:-2, k], min(hk, 0.05 * c.dx ** 2 / c.dt), rtol=1e-12, atol=1e-300)
    if ushear > 1:
        assert K[4, 4, 3] == kmax and KH[4, 4, 3] == 0.05 * c.dx ** 2 / c.dt  # both bounds reached


@pytest.mark.parametrize("backend", BACKENDS)
def test_theta_diffusion_of_a_sinusoid(backend):
    c = _shear_case(1e-2, 0.0)
    i = np.arange(c.nx) - 3
    c.fields["th"] = np.sin(2 * np.pi * i / (c.nx - 6))[:, None, None] * np.ones((1, c.ny, c.nz))
    b = _backend(backend, c)
    KH, T = b.get_turbulence(1), b.get_turbulence(5)
    th = c.fields["th"]
    for k in range(2, c.nz - 3):
        kh = KH[3:-3, 3:-3, k]
        expect = PR_INV * kh * (th[4:-2, 3:-3, k] - 2 * th[3:-3, 3:-3, k] + th[2:-4, 3:-3, k]) / c.dx ** 2
        np.testing.assert_allclose(T[3:-3, 3:-3, k], expect, rtol=1e-10, atol=1e-12 * np.abs(expect).max())


@pytest.mark.parametrize("backend", BACKENDS)
def test_vertical_diffusion_conserves_the_column(backend):
    c = _shear_case(1e-2, 0.0)
    rng = np.random.default_rng(5)
    prof = rng.standard_normal(c.nz)
    c.fields["th"] = np.broadcast_to(prof, (c.nx, c.ny, c.nz)).copy()
    b = _backend(backend, c)
    T = b.get_turbulence(5)
    mu = c.columns["mu"]
    col = (T[5, 5, 1:c.nz - 1] / mu[1:c.nz - 1]).sum() * c.dz   # sum of the flux differences
    assert abs(col) <= 1e-12 * np.abs(T[5, 5]).max() * c.dz * c.nz
    assert np.abs(T[5, 5, 1:c.nz - 1]).max() > 0


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("land", [False, True])
def test_momentum_friction_surface_drag_and_lid(backend, land):
    ushear, u0 = 1e-2, 7.0
    c = _shear_case(ushear, 0.0, u0=u0)
    c.fields["v"][:] = 3.0
    landsea = np.full((c.nx, c.ny), 0.0 if land else 1.0)
    b = _backend(backend, c, landsea)
    K, U, V, W = b.get_turbulence(0), b.get_turbulence(2), b.get_turbulence(3), b.get_turbulence(4)
    z = c.columns["zsu"]
    u1 = u0 + ushear * z[1]
    speed = np.hypot(u1, 3.0)
    cd = 3e-3 if land else 1.1e-3 + 4e-5 * speed
    kin = K[4, 4, 3]
    assert kin > 0
    sl = np.s_[3:-3, 3:-3]
    # lowest level: the resolved stress above minus the surface drag, over the layer
    np.testing.assert_allclose(U[sl][..., 1], (kin * ushear - cd * speed * u1) / c.dz, rtol=1e-10)
    np.testing.assert_allclose(V[sl][..., 1], (0.0 - cd * speed * 3.0) / c.dz, rtol=1e-10)
    # interior: uniform K and linear u, so no net stress divergence
    assert np.abs(U[sl][..., 2:c.nz - 2]).max() <= 1e-12 * kin * ushear / c.dz
    assert np.abs(W[sl][..., 2:c.nz - 1]).max() <= 1e-12 * kin * ushear / c.dz
    # the lid: K = 0 above the top level
    np.testing.assert_allclose(U[sl][..., c.nz - 2], -kin * ushear / c.dz, rtol=1e-10)
