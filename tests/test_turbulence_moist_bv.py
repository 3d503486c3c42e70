"""The saturated branch of the Brunt-Vaisala frequency of the turbulence closure (turbulence.cpp:644-690),
pinned on the oracle.

N^2 is read back from the mixing coefficient: in a uniform vertical shear S the closure gives
K = (c_s l)^2 sqrt(S^2 - N^2 / Pr) (turbulence.cpp:737-831, pinned for the dry branch in
test_turbulence_closed_form.py), so N^2 = Pr (S^2 - (K / (c_s l)^2)^2).  The saturation flags are set
directly (every point saturated), so the states below need not come from a Kessler call.

- Cold, dry limit: where the saturation vapour pressure vanishes (below 145 K) and the air holds no water, the
  saturated branch must give the dry one, g / theta dtheta/dz, with theta at the w-level the mean of its
  two levels.
- Water loading: N^2 of a w-level falls by g/dz for each unit of condensate (cloud or rain) added at its
  upper level and rises by as much for the lower level (the -g dq_w/dz term), exactly.
- The moist-adiabatic factor of Durran and Klemp (1982, J. Atmos. Sci. 39, 2152-2158, their eq. 36),
  F = (1 + L q / (R_d T)) / (1 + eps L^2 q / (c_p R_d T^2)): vapour added uniformly to both levels of a
  saturated w-level changes neither the equivalent potential temperature difference nor the water
  gradient, so the latent-and-stability part of N^2 scales by F(q + dq) / F(q).
- The whole: N^2 = g / theta_w * F * d(theta_e)/dz - g dq_w/dz, with the equivalent potential
  temperature linearised as theta (1 + L q_s / (c_p T)) (the reference's choice) and q_s = eps e_s /
  (p - e_s) from the Goff-Gratch saturation vapour pressure, a different published formula from the
  reference's (Bolton 1980); over a saturated warm column the two agree to 0.5 % of the latent part.
"""
import numpy as np
import pytest

from helpers import zero_case
from oracle.pyoracle import Oracle

# the model's constants (model.h:250-258) and the closure's (turbulence.cpp)
G, LV, RD, CP, P0 = 9.8, 2.5e6, 287.0, 1004.0, 1.0e5
EPS = 0.62197  # R_d / R_v
CS, LVMAX, PR_INV = 0.25, 200.0, 3.0
SHEAR = 0.05


def es_goff_gratch(T):
    """Saturation vapour pressure over liquid water (Pa), Goff and Gratch (1946) as the WMO (1988) gives
    it, with the steam-point temperature 373.15 K and pressure 1013.25 hPa."""
    ts = 373.15
    x = ts / T
    log10_ew = (-7.90298 * (x - 1) + 5.02808 * np.log10(x) - 1.3816e-7 * (10 ** (11.344 * (1 - 1 / x)) - 1)
                + 8.1328e-3 * (10 ** (-3.49149 * (x - 1)) - 1) + np.log10(1013.25))
    return 100.0 * 10 ** log10_ew


def test_goff_gratch_against_tables():
    """The formula as typed here gives the tabulated values (Smithsonian Meteorological Tables): 6.112 hPa
    at 0 C, 12.27 at 10 C, 23.37 at 20 C, 42.43 at 30 C."""
    for tc, hpa in ((0.0, 6.112), (10.0, 12.27), (20.0, 23.37), (30.0, 42.43)):
        assert abs(es_goff_gratch(273.15 + tc) / 100.0 - hpa) <= 2e-3 * hpa, tc


def _column(c, theta_sfc, dthdz, p_sfc):
    """theta rising linearly with height, the Exner function from hydrostatic balance
    (d Pi/dz = -g / (c_p theta)) from p_sfc at z = 0, temperature and pressure at the scalar levels."""
    z = c.columns["zsu"]
    theta = theta_sfc + dthdz * z
    zf = np.linspace(0.0, z.max(), 4001)
    tf = theta_sfc + dthdz * zf
    pi_f = (p_sfc / P0) ** (RD / CP) - G / CP * np.concatenate([[0.0], np.cumsum(0.5 * (1 / tf[1:] + 1 / tf[:-1]) * np.diff(zf))])
    exner = np.interp(z, zf, pi_f)
    below = z < 0  # the ghost level below the ground
    exner[below] = (p_sfc / P0) ** (RD / CP) - G / CP * z[below] / theta_sfc
    return theta, exner, theta * exner, P0 * exner ** (CP / RD)


def _case(theta_sfc, dthdz, p_sfc, vapour, nz=22):
    c = zero_case(nx=10, ny=8, nz=nz, microphysics=1, dz=250.0)
    theta, exner, T, p = _column(c, theta_sfc, dthdz, p_sfc)
    c.columns["tb"] = theta.copy()
    c.columns["tbv"] = theta.copy()
    tbw = theta.copy()
    tbw[1:] = 0.5 * (theta[1:] + theta[:-1])
    c.columns["tbw"] = tbw
    c.columns["qb"] = np.asarray(vapour(T, p), dtype=np.float64)
    c.fields["pbar"][:] = exner[None, None, :]
    c.fields["u"][:] = (SHEAR * c.columns["zsu"])[None, None, :]
    return c, theta, T, p


def _n2(c, saturated=True):
    o = Oracle(c)
    o.set_turbulence(True)
    o.set_saturated(np.full((c.nx, c.ny, c.nz), 1 if saturated else 0))
    o.calculate_diff_tend(1, c.nx - 1, 1, c.ny - 1)
    K = o.get_turbulence(0)[c.nx // 2, c.ny // 2]
    l2 = min(c.dz, LVMAX) ** 2
    ks = range(2, c.nz - 1)
    assert all(0 < K[k] < 0.15 * l2 / c.dt for k in ks)
    return np.array([(SHEAR ** 2 - (K[k] / (CS ** 2 * l2)) ** 2) / PR_INV for k in ks])


def _qs(T, p):
    e = es_goff_gratch(T)
    return EPS * e / (p - e)


def _factor(q, T):
    return (1 + LV * q / (RD * T)) / (1 + 0.622 * LV * LV * q / (CP * RD * T * T))


def test_cold_dry_limit_is_the_dry_branch():
    c, theta, T, _ = _case(200.0, 4e-3, 3.0e4, lambda T, p: np.zeros_like(T))
    assert T.max() < 145.0
    sat, dry = _n2(c, True), _n2(c, False)
    np.testing.assert_allclose(sat, dry, rtol=1e-7)
    ks = np.arange(2, c.nz - 1)
    expect = G / c.columns["tbw"][ks] * (theta[ks] - theta[ks - 1]) / c.dz
    np.testing.assert_allclose(sat, expect, rtol=1e-7)


@pytest.mark.parametrize("field", ["qc", "qr"])
def test_water_loading(field):
    c, _, _, _ = _case(295.0, 3e-3, 1.0e5, _qs)
    base = _n2(c)
    k, dq = 9, 1e-3
    c.fields[field][:, :, k] += dq
    loaded = _n2(c)
    d = loaded - base  # entry m is the w-level m + 2
    np.testing.assert_allclose(d[k - 2], -G * dq / c.dz, rtol=1e-9)
    np.testing.assert_allclose(d[k + 1 - 2], G * dq / c.dz, rtol=1e-9)
    others = [m for m in range(len(d)) if m not in (k - 2, k - 1)]
    assert np.abs(d[others]).max() <= 1e-12 * np.abs(base).max()


def test_durran_klemp_factor():
    c, _, T, _ = _case(295.0, 3e-3, 1.0e5, _qs)
    ks = np.arange(2, c.nz - 1)
    qw = c.columns["qb"]
    loading = G * (qw[ks] - qw[ks - 1]) / c.dz
    base = _n2(c) + loading
    dq = 4e-3
    c.columns["qb"] = qw + dq
    more = _n2(c) + loading  # the water gradient is unchanged
    qm, Tm = 0.5 * (qw[ks] + qw[ks - 1]), 0.5 * (T[ks] + T[ks - 1])
    np.testing.assert_allclose(more / base, _factor(qm + dq, Tm) / _factor(qm, Tm), rtol=1e-10)


@pytest.mark.parametrize("theta_sfc,dthdz", [(298.0, 3e-3), (290.0, 6e-3), (302.0, 1e-3)])
def test_saturated_column_against_published_forms(theta_sfc, dthdz):
    c, theta, T, p = _case(theta_sfc, dthdz, 1.0e5, _qs)
    ks = np.arange(2, c.nz - 1)
    qs = c.columns["qb"]
    latent = theta * LV * qs / (CP * T)  # theta_e - theta, linearised
    Tm, qm = 0.5 * (T[ks] + T[ks - 1]), 0.5 * (qs[ks] + qs[ks - 1])
    scale = G / c.columns["tbw"][ks] * _factor(qm, Tm) / c.dz
    expect = scale * ((theta + latent)[ks] - (theta + latent)[ks - 1]) - G * (qs[ks] - qs[ks - 1]) / c.dz
    got = _n2(c)
    tol = 5e-3 * np.abs(scale * (latent[ks] - latent[ks - 1]))
    assert np.all(np.abs(got - expect) <= tol), np.abs(got - expect) / tol
    # a latent term 2 % larger (a saturation vapour pressure 2 % off) would fail
    wrong = scale * ((theta + 1.02 * latent)[ks] - (theta + 1.02 * latent)[ks - 1]) - G * (qs[ks] - qs[ks - 1]) / c.dz
    assert np.any(np.abs(got - wrong) > tol)
