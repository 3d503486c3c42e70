"""Pins of the oracle's building blocks against mathematics (no GPU).

FFT against numpy.fft; face interpolants against polynomial exactness and the closed form of their
upwind term; the Poisson solve against a manufactured solution; Kessler against its conservation laws.
"""
import ctypes

import numpy as np
import pytest

from model_amd import configs, ids
from oracle.pyoracle import Oracle, _ptr


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6, 7, 8, 12, 14, 19, 22, 26, 38, 46, 60, 64, 94, 96, 128, 160, 184, 190, 192, 194,
                               256, 512, 777, 1029])
@pytest.mark.parametrize("sign", [-1, 1])
def test_fft1d_matches_numpy(oracle_lib, n, sign):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    buf = np.empty(2 * n)
    buf[0::2], buf[1::2] = z.real, z.imag
    oracle_lib.orc_fft1d(n, _ptr(buf), sign)
    got = buf[0::2] + 1j * buf[1::2]
    want = np.fft.fft(z) if sign < 0 else np.fft.ifft(z) * n  # unnormalised backward, FFTW convention
    assert np.abs(got - want).max() <= 1e-13 * max(1.0, np.abs(want).max())


@pytest.mark.parametrize("shape", [(64, 64), (96, 128), (160, 192), (256, 256), (194, 100), (184, 190)])
def test_fft2d_matches_numpy(oracle_lib, shape):
    n0, n1 = shape
    rng = np.random.default_rng(7)
    z = rng.standard_normal((3, n0, n1)) + 1j * rng.standard_normal((3, n0, n1))
    buf = np.empty((3, n0, n1, 2))
    buf[..., 0], buf[..., 1] = z.real, z.imag
    buf = np.ascontiguousarray(buf)
    oracle_lib.orc_fft2d(n0, n1, 3, _ptr(buf), -1)
    got = buf[..., 0] + 1j * buf[..., 1]
    want = np.fft.fft2(z, axes=(1, 2))
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def _face_value(fn, poly, sign):
    # the flux-form interpolants act on cell averages: cells centred at x = -2..3 (unit width),
    # face between the cells at x=0 and x=1
    x = np.arange(-2, 4, dtype=np.float64)
    anti = poly.integ()
    p = np.ascontiguousarray(anti(x + 0.5) - anti(x - 0.5))
    arr = (ctypes.c_double * len(p))(*p)
    ptr = ctypes.cast(ctypes.byref(arr, 2 * 8), ctypes.POINTER(ctypes.c_double))
    return fn(ptr, sign) if sign is not None else fn(ptr)


@pytest.mark.parametrize("deg", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_interp5_exact_up_to_degree_4(oracle_lib, deg, sign):
    rng = np.random.default_rng(deg)
    c = rng.standard_normal(deg + 1)
    poly = np.polynomial.Polynomial(c)
    v = _face_value(oracle_lib.orc_interp5, poly, sign)
    assert v == pytest.approx(poly(0.5), rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_interp5_upwind_term_closed_form(oracle_lib, sign):
    # x^5: the 6th-order centred part is exact; the upwind part is -sign/60 * (5th difference = 5!),
    # i.e. the textbook 5th-order upwind face value (2,-13,47,27,-3)/60 for a positive velocity
    poly = np.polynomial.Polynomial([0, 0, 0, 0, 0, 1.0])
    v = _face_value(oracle_lib.orc_interp5, poly, sign)
    assert v == pytest.approx(0.5 ** 5 - 2.0 * sign, rel=1e-13)
    x = np.arange(-2, 4, dtype=np.float64)
    avg = poly.integ()(x + 0.5) - poly.integ()(x - 0.5)
    w = np.array([2, -13, 47, 27, -3, 0]) / 60.0 if sign > 0 else np.array([0, -3, 27, 47, -13, 2]) / 60.0
    assert v == pytest.approx(float(w @ avg), rel=1e-12)


@pytest.mark.parametrize("deg", [0, 1, 2])
@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_interp3_exact_up_to_degree_2(oracle_lib, deg, sign):
    c = np.random.default_rng(10 + deg).standard_normal(deg + 1)
    poly = np.polynomial.Polynomial(c)
    v = _face_value(oracle_lib.orc_interp3, poly, sign)
    assert v == pytest.approx(poly(0.5), rel=1e-13, abs=1e-13)


@pytest.mark.parametrize("sign", [-1.0, 1.0])
def test_interp3_upwind_term_closed_form(oracle_lib, sign):
    # x^3: the 4th-order centred part is exact; the upwind part adds sign/12 * (3rd difference = 3!),
    # i.e. the textbook 3rd-order upwind face value (-1,5,2)/6 for a positive velocity
    poly = np.polynomial.Polynomial([0, 0, 0, 1.0])
    v = _face_value(oracle_lib.orc_interp3, poly, sign)
    assert v == pytest.approx(0.125 + 0.5 * sign, rel=1e-13)
    x = np.arange(-2, 4, dtype=np.float64)
    avg = poly.integ()(x + 0.5) - poly.integ()(x - 0.5)
    w = np.array([0, -1, 5, 2, 0, 0]) / 6.0 if sign > 0 else np.array([0, 0, 2, 5, -1, 0]) / 6.0
    assert v == pytest.approx(float(w @ avg), rel=1e-12)


def test_interp2_is_the_mean(oracle_lib):
    poly = np.polynomial.Polynomial([3.0, -2.0])  # linear: cell averages are point values
    assert _face_value(oracle_lib.orc_interp2, poly, None) == pytest.approx(poly(0.5))


def _discrete_laplacian(case, pi):
    """The operator whose Fourier diagonalisation is init_fft3d's tridiagonal system
    (pressure.cpp:779-792): rhou_k (d2x + d2y)/dx^2 + a_k (p_{k-1}-p_k) + c_k (p_{k+1}-p_k), a_1 = c_{nk-2} = 0."""
    col = case.columns
    nz, dx, dz = case.nz, case.dx, case.dz
    mu, mw, rhow, rhou = col["mu"], col["mw"], col["rhow"], col["rhou"]
    out = np.zeros_like(pi)
    for k in range(1, nz - 1):
        p = pi[k]
        lap = (np.roll(p, 1, 0) + np.roll(p, -1, 0) + np.roll(p, 1, 1) + np.roll(p, -1, 1) - 4 * p) / dx ** 2
        v = np.zeros_like(p)
        if k > 1:
            v += mu[k] * mw[k] * rhow[k] / dz ** 2 * (pi[k - 1] - p)
        if k < nz - 2:
            v += mu[k] * mw[k + 1] * rhow[k + 1] / dz ** 2 * (pi[k + 1] - p)
        out[k] = rhou[k] * lap + v
    return out


@pytest.mark.parametrize("dims", [(16, 16, 12), (24, 20, 10), (64, 64, 40)])
def test_poisson_manufactured_solution(dims):
    nx, ny, nz = dims
    case = configs.small_case(nx=nx, ny=ny, nz=nz)
    o = Oracle(case)
    rng = np.random.default_rng(3)
    pi = np.zeros((nz, nx, ny))
    pi[1:nz - 1] = rng.standard_normal((nz - 2, nx, ny))
    rhs = np.ascontiguousarray(_discrete_laplacian(case, pi))
    x = np.zeros_like(pi)
    o.L.orc_poisson_rhs(o.h, _ptr(rhs), _ptr(x))
    # gauge: the horizontal-mean mode is pinned to zero at k = nz-2 (pressure.cpp:858-859)
    want = pi - pi[nz - 2].mean()
    err = np.abs(x[1:nz - 1] - want[1:nz - 1]).max()
    assert err <= 1e-10 * np.abs(want).max()


def test_kessler_point_conservation_laws(oracle_lib):
    rng = np.random.default_rng(11)
    case = configs.supercell(nx=16, ny=16, nz=40, bubble=False)
    col = case.columns
    n_checked = 0
    for trial in range(4000):
        k = int(rng.integers(1, case.nz - 1))
        th = ctypes.c_double(rng.normal(0, 3))
        qv = ctypes.c_double(rng.normal(0, 2e-3))
        qc = ctypes.c_double(abs(rng.normal(0, 2e-3)) * (rng.random() < 0.7))
        qr = ctypes.c_double(abs(rng.normal(0, 2e-3)) * (rng.random() < 0.7))
        pi = rng.normal(0, 50)
        pbar = col["pib"][k]
        th0, qv0, qc0, qr0 = th.value, qv.value, qc.value, qr.value
        err = oracle_lib.orc_kessler_point(case.dt, col["rhou"][k], col["tb"][k], col["qb"][k], col["tbv"][k], 0.0, 0.0, pbar,
                                           pi, ctypes.byref(th), ctypes.byref(qv), ctypes.byref(qc), ctypes.byref(qr))
        assert err == 0
        # total water is conserved point-wise
        assert qv.value + qc.value + qr.value == pytest.approx(qv0 + qc0 + qr0, abs=1e-15)
        # cloud and rain stay non-negative
        assert qc.value >= 0 and qr.value >= 0
        # the heating matches the latent heat of the net condensation: dtheta = Lv/(cp*pi_full) * (-dqv)
        pfull = pi / (1004.0 * col["tbv"][k]) + pbar
        assert th.value - th0 == pytest.approx(2.5e6 / (1004.0 * pfull) * (qv0 - qv.value), rel=1e-9, abs=1e-12)
        n_checked += 1
    assert n_checked == 4000


def test_kessler_saturation_adjustment_reaches_saturation(oracle_lib):
    # a supersaturated parcel without rain ends near saturation (first-order latent-heat correction)
    case = configs.supercell(nx=16, ny=16, nz=40, bubble=False)
    col = case.columns
    k = 6
    th, qv, qc, qr = (ctypes.c_double(v) for v in (0.0, 3e-3, 0.0, 0.0))
    oracle_lib.orc_kessler_point(case.dt, col["rhou"][k], col["tb"][k], col["qb"][k], col["tbv"][k], 0.0, 0.0, col["pib"][k],
                                 0.0, ctypes.byref(th), ctypes.byref(qv), ctypes.byref(qc), ctypes.byref(qr))
    temp = (th.value + col["tb"][k]) * col["pib"][k]
    pd = 1e5 * col["pib"][k] ** (1004.0 / 287.0)
    qsat = configs.qv_sat(temp, pd)
    assert qc.value > 0
    assert abs(qv.value + col["qb"][k] - qsat) < 0.05 * (3e-3)


def test_kessler_thresholds_inactive_branches(oracle_lib):
    # qc below 1e-3 and qr below 1e-12: no autoconversion, accretion or evaporation; a subsaturated
    # parcel without cloud is left untouched
    case = configs.supercell(nx=16, ny=16, nz=40, bubble=False)
    col = case.columns
    k = 20
    th, qv, qc, qr = (ctypes.c_double(v) for v in (0.3, -1e-3, 0.0, 1e-13))
    oracle_lib.orc_kessler_point(case.dt, col["rhou"][k], col["tb"][k], col["qb"][k], col["tbv"][k], 0.0, 0.0, col["pib"][k],
                                 0.0, ctypes.byref(th), ctypes.byref(qv), ctypes.byref(qc), ctypes.byref(qr))
    assert (th.value, qv.value, qc.value, qr.value) == (0.3, -1e-3, 0.0, 1e-13)


def test_fall_speed_closed_form(oracle_lib):
    assert oracle_lib.orc_fall_speed(1e-12, 1.1, 1.1) == 0.0
    assert oracle_lib.orc_fall_speed(-1.0, 1.1, 1.1) == 0.0
    # 36.34 * (1e-3 * rho * qr)^0.1364 * sqrt(rho1 / rho)
    v = oracle_lib.orc_fall_speed(2e-3, 0.5, 1.0)
    assert v == pytest.approx(36.34 * (1e-3 * 0.5 * 2e-3) ** 0.1364 * np.sqrt(2.0), rel=1e-14)


def test_base_state_hydrostatic_and_density():
    case = configs.warm_bubble()
    col = case.columns
    # rhow is the arithmetic mean of the adjacent rhou (initializer.cpp:905)
    assert np.allclose(col["rhow"][2:], 0.5 * (col["rhou"][2:] + col["rhou"][1:-1]), rtol=1e-15)
    # Exner pressure decreases with height at the hydrostatic rate g / (cp theta_v)
    dpi = np.diff(col["pib"][1:-1])
    dz = np.diff(col["zsu"][1:-1])
    tv = 0.5 * (col["tbv"][2:-1] + col["tbv"][1:-2])
    assert np.allclose(dpi, -9.8 * dz / (1004.0 * tv), rtol=1e-12)
    # unstretched grid: mu = mw = 1 when the lowest level sits at dz
    assert np.allclose(col["mu"], 1.0) and np.allclose(col["mw"], 1.0)


def test_kessler_guard_flags_the_oracle_state():
    """kessler.cpp:83-88: E > 100 is the reference's fatal condition; the oracle records it (orc_error),
    for the same forced state the GPU test feeds the HIP library (an absurd rain mixing ratio)."""
    import numpy as np
    from helpers import zero_case
    from model_amd import ids
    from oracle.pyoracle import Oracle
    for q, flagged in ((1e6, 0), (1e8, 1)):
        c = zero_case(nx=16, ny=16, nz=12, microphysics=1)
        o = Oracle(c)
        qr = np.zeros((c.nx, c.ny, c.nz))
        qr[8, 8, 3] = q
        o.set_field(ids.QR, ids.ROLE_P, qr)
        o.set_field(ids.QV, ids.ROLE_P, np.full((c.nx, c.ny, c.nz), -0.5 * c.columns["qb"][3]))
        o.kessler(1, c.nx - 1, 1, c.ny - 1)
        assert o.error == flagged
