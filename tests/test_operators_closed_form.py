"""Closed-form checks of each reference operator, run on the oracle (CPU) and on the HIP library (GPU).

Every expected value below follows from the mathematics of the scheme on a specially built state
(linear fields, uniform winds, unit density), not from the code under test.
"""
import numpy as np
import pytest

from helpers import BACKENDS, index_grids, make_backend, zero_case
from model_amd import ids

P = ids.ROLE_P


@pytest.mark.parametrize("backend", BACKENDS)
def test_u_advection_by_v_uses_the_staggered_advecting_velocity(backend):
    # U linear in j, V linear in i: d(vu)/dy = va * dU/dy with va at x_{i-1/2}
    c = zero_case(unit_density=True)
    I, J, K = index_grids(c)
    a, b, cc, d = 0.3, 1.0, 0.2, -0.7
    c.fields["u"] = a * J + b
    c.fields["v"] = cc * I + d
    m = make_backend(backend, c)
    step = 0.5
    m.advect_uvw(step, 3, c.nx - 3, 3, c.ny - 3)
    up = m.get_field(ids.U, P)
    va = cc * (I - 0.5) + d
    want = c.fields["u"] - step * (c.dt / c.dx) * va * a
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(up[sl] - want[sl]).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_v_advection_by_u_uses_the_staggered_advecting_velocity(backend):
    c = zero_case(unit_density=True)
    I, J, K = index_grids(c)
    a, b, cc, d = 0.25, -2.0, -0.15, 0.5
    c.fields["v"] = a * I + b
    c.fields["u"] = cc * J + d
    m = make_backend(backend, c)
    step = 1.0
    m.advect_uvw(step, 3, c.nx - 3, 3, c.ny - 3)
    vp = m.get_field(ids.V, P)
    ua = cc * (J - 0.5) + d
    want = c.fields["v"] - step * (c.dt / c.dx) * ua * a
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(vp[sl] - want[sl]).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_u_vertical_flux_and_ground(backend):
    # U linear in k, uniform W, unit density: -dtz*w*dU/dk inside, and the k=0 flux is zero
    c = zero_case(unit_density=True)
    I, J, K = index_grids(c)
    a, b, w0 = 0.4, 1.5, 0.8
    c.fields["u"] = a * K + b
    c.fields["w"] = np.full_like(K, w0)
    m = make_backend(backend, c)
    step = 1.0 / 3.0
    m.advect_uvw(step, 3, c.nx - 3, 3, c.ny - 3)
    up = m.get_field(ids.U, P)
    dtz = c.dt / c.dz
    want = c.fields["u"] - step * dtz * w0 * a
    want[:, :, 1] = c.fields["u"][:, :, 1] - step * dtz * w0 * (a * 1.5 + b)
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(up[sl] - want[sl]).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
@pytest.mark.parametrize("moist", [0, 1])
def test_buoyancy_overwrites_w(backend, moist):
    c = zero_case(microphysics=moist)
    I, J, K = index_grids(c)
    theta0 = 0.7
    c.fields["th"] = theta0 * c.columns["tbv"][None, None, :] * np.ones_like(K)
    c.fields["w"] = 3.0 * np.sin(I + 2 * J + K)  # any w flux tendency must be discarded
    c.fields["u"] = np.cos(I - J) * 2.0
    b = theta0
    if moist:
        c.fields["qv"] = np.full_like(K, 1e-3)
        c.fields["qc"] = np.full_like(K, 2e-4)
        c.fields["qr"] = np.full_like(K, 1e-4)
        b = theta0 + 0.61 * 1e-3 - 2e-4 - 1e-4
    m = make_backend(backend, c)
    step = 0.5
    m.advect_uvw(step, 3, c.nx - 3, 3, c.ny - 3)
    wp = m.get_field(ids.W, P)
    want = c.fields["w"] + step * c.dt * 9.8 * b
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 2:c.nz - 1]
    assert np.abs(wp[sl] - want[sl]).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_theta_horizontal_flux_perturbation_and_basic_state(backend):
    c = zero_case(unit_density=True)
    I, J, K = index_grids(c)
    u0, a, bb = 1.3, 0.05, -0.02
    c.fields["u"] = np.full_like(I, u0)
    c.fields["th"] = a * I + 1.0
    c.fields["thbar"] = bb * I + 2.0
    c.columns["tb"] = np.full(c.nz, 300.0)
    m = make_backend(backend, c)
    step = 1.0
    m.advect_theta(step, 3, c.nx - 3, 3, c.ny - 3)
    thp = m.get_field(ids.TH, P)
    # flux = u*(thbar_face) + (ubar+u)*(th_face): divergence u0*(bb + a)
    want = c.fields["th"] - step * (c.dt / c.dx) * u0 * (a + bb)
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(thp[sl] - want[sl]).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_theta_base_state_advection(backend):
    # tb linear in k, uniform w, unit density: -dtz * w * dtb/dk (Z_ADVECT_BASE, advection.cpp:16-18)
    c = zero_case(unit_density=True)
    I, J, K = index_grids(c)
    s, w0 = 0.9, -0.6
    c.columns["tb"] = 300.0 + s * np.arange(c.nz)
    c.fields["w"] = np.full_like(K, w0)
    m = make_backend(backend, c)
    step = 0.5
    m.advect_theta(step, 3, c.nx - 3, 3, c.ny - 3)
    thp = m.get_field(ids.TH, P)
    want = -step * (c.dt / c.dz) * w0 * s
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(thp[sl] - want).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_theta_flux_form_conserves_mass_weighted_heat(backend):
    # compact theta' away from every boundary, arbitrary winds: sum(rhou * dtheta) telescopes to zero
    c = zero_case()
    I, J, K = index_grids(c)
    rng = np.random.default_rng(5)
    c.fields["u"] = rng.standard_normal(I.shape) * 4
    c.fields["v"] = rng.standard_normal(I.shape) * 4
    c.fields["w"] = rng.standard_normal(I.shape) * 2
    c.fields["w"][:, :, c.nz - 1] = 0.0
    c.fields["ubar"] = 5 + rng.standard_normal(I.shape)
    c.fields["vbar"] = rng.standard_normal(I.shape)
    c.columns["tb"] = np.full(c.nz, 300.0)
    th = np.zeros(I.shape)
    th[8:14, 6:10, 3:8] = rng.standard_normal((6, 4, 5))  # three cells clear of the summed region
    c.fields["th"] = th
    m = make_backend(backend, c)
    m.advect_theta(1.0, 3, c.nx - 3, 3, c.ny - 3)
    thp = m.get_field(ids.TH, P)
    rhou = c.columns["rhou"]
    tot = np.sum((thp - th)[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1] * rhou[None, None, 1:c.nz - 1])
    scale = np.sum(np.abs(thp - th)[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1] * rhou[None, None, 1:c.nz - 1])
    assert scale > 0 and abs(tot) < 1e-12 * scale


@pytest.mark.parametrize("backend", BACKENDS)
def test_rain_sedimentation_closed_form(backend):
    # uniform rain, no wind, unit density: the fall-speed flux is uniform except below the lid
    c = zero_case(microphysics=1, unit_density=True)
    q0 = 2e-3
    c.fields["qr"] = np.full((c.nx, c.ny, c.nz), q0)
    m = make_backend(backend, c)
    step = 1.0
    m.advect_moisture(step, 3, c.nx - 3, 3, c.ny - 3)
    qrp = m.get_field(ids.QR, P)
    vt = 36.34 * (q0 * 0.001) ** 0.1364
    want = np.full(qrp.shape, q0)
    want[:, :, c.nz - 2] = q0 - step * (c.dt / c.dz) * 0.5 * vt * q0  # level NZ-1 keeps a zero fall speed
    sl = np.s_[3:c.nx - 3, 3:c.ny - 3, 1:c.nz - 1]
    assert np.abs(qrp[sl] - want[sl]).max() < 1e-15
    vtf = m.get_field(ids.VT, 0)
    assert np.abs(vtf[3:c.nx - 3, 3:c.ny - 3, 0:c.nz - 1] - vt).max() < 1e-12


@pytest.mark.parametrize("backend", BACKENDS)
def test_pressure_gradient_force_closed_form(backend):
    c = zero_case()
    I, J, K = index_grids(c)
    a, b, d = 3.0, -2.0, 5.0
    c.fields["pi"] = a * I + b * J + d * K
    m = make_backend(backend, c)
    step = 0.5
    m.pgf(step, 1, c.nx - 1, 1, c.ny - 1)
    up, vp, wp = (m.get_field(f, P) for f in (ids.U, ids.V, ids.W))
    dtx, dtz = c.dt / c.dx, c.dt / c.dz
    assert np.abs(up[1:-1, 1:-1, 1:-1] + step * dtx * a).max() < 1e-12
    assert np.abs(vp[1:-1, 1:-1, 1:-1] + step * dtx * b).max() < 1e-12
    assert np.abs(wp[1:-1, 1:-1, 2:-1] + step * dtz * c.columns["mw"][None, None, 2:-1] * d).max() < 1e-12
    assert np.all(wp[:, :, 1] == 0) and np.all(up[0] == 0) and np.all(up[-1] == 0)


def _divergence(c, up, vp, wp):
    col = c.columns
    k = np.arange(1, c.nz - 1)
    d = (col["rhou"][k] * (up[4:c.nx - 2, 1:c.ny - 1][:, :, k] - up[3:c.nx - 3, 1:c.ny - 1][:, :, k]) / c.dx +
         col["rhou"][k] * (vp[3:c.nx - 3, 2:c.ny][:, :, k] - vp[3:c.nx - 3, 1:c.ny - 1][:, :, k]) / c.dx +
         (col["rhow"][k + 1] * wp[3:c.nx - 3, 1:c.ny - 1][:, :, k + 1] - col["rhow"][k] * wp[3:c.nx - 3, 1:c.ny - 1][:, :, k]) *
         col["mu"][k] / c.dz)
    return d  # [i-3, j-1, k-1]


@pytest.mark.parametrize("backend", BACKENDS)
def test_projection_removes_the_divergence(backend):
    """solve_pressure followed by the pressure-gradient force leaves a divergence-free flow wherever the
    FFT operator and the gradient stencil coincide (rows 1..NY-3); the mean mode keeps the column
    integral of the divergence at the top level (the singular row the reference leaves unsolved)."""
    c = zero_case(nx=16, ny=16, nz=12)
    rng = np.random.default_rng(9)
    shape = (c.nx, c.ny, c.nz)
    for f in ("u", "v", "w"):
        c.fields[f] = rng.standard_normal(shape)
    c.fields["w"][:, :, 0:2] = 0.0
    c.fields["w"][:, :, c.nz - 1] = 0.0
    m = make_backend(backend, c)
    step = 0.5
    m.solve_pressure(step)
    m.pgf(step, 1, c.nx - 1, 1, c.ny - 1)
    up, vp, wp = (m.get_field(f, P) for f in (ids.U, ids.V, ids.W))
    d0 = _divergence(c, *(c.fields[f] for f in ("u", "v", "w")))
    d = _divergence(c, up, vp, wp)
    rows = np.s_[:, 0:c.ny - 3, :]
    inner = d[rows].copy()
    top = inner[:, :, -1]
    inner[:, :, -1] = top - top.mean()
    assert np.abs(inner).max() < 1e-10 * np.abs(d0).max()


@pytest.mark.parametrize("backend", BACKENDS)
def test_rayleigh_profile(backend):
    c = zero_case(nx=12, ny=10, nz=14, rayleigh_level=8)
    ones = np.ones((c.nx, c.ny, c.nz))
    for f in ("u", "v", "w", "th"):
        c.fields[f] = ones.copy()
    m = make_backend(backend, c)
    m.rayleigh(1, c.nx - 1, 1, c.ny - 1, 8)
    up, wp = m.get_field(ids.U, P), m.get_field(ids.W, P)
    zu, zw = c.columns["zu"], c.columns["zw"]
    k = np.arange(8, c.nz)
    cu = 0.05 * 0.5 * (1 - np.cos(np.pi * (zu[k] - zu[8]) / (zu[-1] - zu[8])))
    cw = 0.05 * 0.5 * (1 - np.cos(np.pi * (zw[k] - zu[8]) / (zw[-1] - zu[8])))
    assert np.allclose(up[1:-1, 1:-1, 8:], 1 - cu[None, None, :], rtol=0, atol=1e-15)
    assert np.allclose(wp[1:-1, 1:-1, 8:], 1 - cw[None, None, :], rtol=0, atol=1e-15)
    assert up[5, 5, -1] == pytest.approx(0.95) and np.all(up[:, :, :8] == 1) and np.all(up[0] == 1)


@pytest.mark.parametrize("backend", BACKENDS)
def test_boundary_conditions_sponge_periodic_lid(backend):
    c = zero_case(nx=12, ny=16, nz=10)
    rng = np.random.default_rng(2)
    for f in ("u", "v", "w", "th", "pi"):
        c.fields[f] = rng.standard_normal((c.nx, c.ny, c.nz))
    m = make_backend(backend, c)
    m.apply_bc()
    u, w, pi = m.get_field(ids.U, P), m.get_field(ids.W, P), m.get_field(ids.PI, 0)
    NX, NY, NZ = c.nx, c.ny, c.nz
    # sponge coefficient 1 on the outermost rows (south j=0, north j=NY-1), 0 from row 5 inwards
    assert np.all(u[1:, 0, 1:NZ - 1] == 0) and np.all(u[1:, NY - 1, 1:NZ - 1] == 0)
    assert np.array_equal(u[3:NX - 3, 5:NY - 5, 1:NZ - 1], c.fields["u"][3:NX - 3, 5:NY - 5, 1:NZ - 1])
    # periodic halos; column i=0 is copied before the lid copy, which starts at i=1, so its ghost
    # levels keep the pre-lid values of column NX-6 (boundaries.cpp:168-175, 304-312)
    assert np.array_equal(u[0, :, 1:NZ - 1], u[NX - 6, :, 1:NZ - 1]) and not np.array_equal(u[0, 1:, 0], u[NX - 6, 1:, 0])
    assert np.array_equal(u[NX - 1], u[5]) and np.array_equal(pi[1], pi[NX - 5])
    # lid: copies for u, zeros for w (from i=1, j=1)
    assert np.array_equal(u[1:, 1:, 0], u[1:, 1:, 1]) and np.array_equal(u[1:, 1:, NZ - 1], u[1:, 1:, NZ - 2])
    assert np.all(w[1:, 1:, [0, 1, NZ - 1]] == 0)


@pytest.mark.parametrize("backend", BACKENDS)
def test_state_at_rest_stays_at_rest(backend):
    c = zero_case(nx=16, ny=16, nz=12, microphysics=1)
    c.columns["qb"] = np.full(c.nz, 1e-3)
    m = make_backend(backend, c)
    m.step(3)
    for f in (ids.U, ids.V, ids.W, ids.TH, ids.QC, ids.QR):
        assert np.abs(m.get_field(f, ids.ROLE_CUR)).max() == 0.0
