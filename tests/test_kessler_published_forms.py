"""Kessler warm rain (kessler.cpp:41-173, 222-252) pinned term by term against published forms.

The oracle's Kessler map is a statement-by-statement restatement of the reference; this file checks
each of its rate terms against the published scheme it implements, not against the oracle's own
expressions:

  * Klemp & Wilhelmson (1978, J. Atmos. Sci. 35:1070) warm-rain rates, in their cgs units (rho in
    g cm^-3, p in mb): autoconversion k1 (qc - a), k1 = 1e-3 s^-1, a = 1e-3; accretion k2 qc qr^0.875,
    k2 = 2.2 s^-1; rain evaporation
        E = (1/rho) (1 - qv/qvs) C (rho qr)^0.525 / (5.4e5 + 2.55e6 / (p qvs)),
        C = 1.6 + 124.9 (rho qr)^0.2046  (ventilation);
    terminal fall speed V = 36.34 (rho qr)^0.1364 (rho_0/rho)^(1/2) m s^-1.
  * Bolton (1980, Mon. Wea. Rev. 108:1046) saturation vapour pressure, checked here against the
    IAPWS saturation pressures of liquid water, and qvs = eps e_s/(p - e_s), eps = M_w/M_d.
  * Saturation adjustment as one Newton step of qv - C = qvs(T + (Lv/cp) C).

The unit conversions are derived in the test (the reference works in SI: rho in kg m^-3, p in Pa),
which explains the reference's 30.3922, 2.03e4, 9.58e6 and its 0.001 factors.  Each check uses inputs
for which the other terms are inactive (no cloud, no rain, or supersaturated so that the evaporation
clamps to zero), and the regimes are chosen so that a 1 % change in any constant fails
(test_evaporation_cases_resolve_every_constant shows it for the evaporation constants).

One deviation from the published scheme is the reference's own: its accretion carries an extra factor
of the base-state density rhou[k] (kessler.cpp:59-63).  The oracle follows the reference; the test
states the factor explicitly.

Both backends run: the oracle (CPU) and the HIP library (-m gpu), through their grid entry points.
"""
import numpy as np
import pytest

from helpers import BACKENDS, make_backend, zero_case
from model_amd import ids

DT = 2.0
P0, RD, CP_REF = 1.0e5, 287.0, 1004.0     # Exner function definition of the model (model.h:250-259)
EPS = 18.01528 / 28.9645                    # molar mass ratio water / dry air
LV_PUB, CP_PUB = 2.501e6, 1005.0            # latent heat at 0 C, dry-air heat capacity
IAPWS_ES = {0.01: 611.657, 10.0: 1228.2, 20.0: 2339.3, 30.0: 4247.0, 40.0: 7384.9}  # Pa, liquid water

# Klemp & Wilhelmson (1978), cgs
K1, A_AUTO, K2, ACC_EXP = 1.0e-3, 1.0e-3, 2.2, 0.875
VENT0, VENT1, VENT_EXP, EVAP_EXP, EVAP_D1, EVAP_D2 = 1.6, 124.9, 0.2046, 0.525, 5.4e5, 2.55e6
VT_C, VT_EXP = 36.34, 0.1364
G_PER_CM3 = 1.0e-3     # kg m^-3 -> g cm^-3
MB_PER_PA = 1.0e-2


class KesslerGrid:
    """Evaluates the Kessler point map on a backend for up to 20 x 14 x 10 points at once.

    Columns: tb = qb = 0, tbv = 300 and pi = 0, so the full potential temperature is THBAR, the Exner
    function is PBAR and the full vapour is QV; rhou (= rhow) is one value for the whole grid."""

    def __init__(self, kind, rho=1.1):
        c = zero_case(nx=16, ny=16, nz=12, microphysics=1, dt=DT)
        for n in ("tb", "qb"):
            c.columns[n] = np.zeros(c.nz)
        c.columns["tbv"] = np.full(c.nz, 300.0)
        for n in ("rhou", "rhow"):
            c.columns[n] = np.full(c.nz, rho)
        for n in ("one_d_rhou", "one_d_rhow"):
            c.columns[n] = np.full(c.nz, 1.0 / rho)
        self.case, self.kind, self.rho = c, kind, rho
        self.b = make_backend(kind, c)
        self.shape = (c.nx - 2, c.ny - 2, c.nz - 2)
        self.capacity = int(np.prod(self.shape))

    def _put(self, vals):
        full = np.full(self.case.nx * self.case.ny * self.case.nz, 0.0).reshape(self.case.nx, self.case.ny, self.case.nz)
        flat = np.empty(self.capacity)
        flat[:vals.size] = vals
        flat[vals.size:] = vals[0]
        full[1:-1, 1:-1, 1:-1] = flat.reshape(self.shape)
        return full

    def run(self, theta, exner, qv, qc, qr):
        theta, exner, qv, qc, qr = np.broadcast_arrays(*[np.atleast_1d(np.asarray(a, dtype=np.float64))
                                                          for a in (theta, exner, qv, qc, qr)])
        n = theta.size
        assert n <= self.capacity
        b = self.b
        b.set_field(ids.THBAR, 0, self._put(theta.ravel()))
        b.set_field(ids.PBAR, 0, self._put(exner.ravel()))
        b.set_field(ids.QBAR, 0, np.zeros((self.case.nx, self.case.ny, self.case.nz)))
        b.set_field(ids.PI, 0, np.zeros((self.case.nx, self.case.ny, self.case.nz)))
        b.set_field(ids.TH, ids.ROLE_P, np.zeros((self.case.nx, self.case.ny, self.case.nz)))
        b.set_field(ids.QV, ids.ROLE_P, self._put(qv.ravel()))
        b.set_field(ids.QC, ids.ROLE_P, self._put(qc.ravel()))
        b.set_field(ids.QR, ids.ROLE_P, self._put(qr.ravel()))
        b.kessler(1, self.case.nx - 1, 1, self.case.ny - 1)
        if self.kind == "hip":
            b.sync()
        out = [b.get_field(f, ids.ROLE_P)[1:-1, 1:-1, 1:-1].ravel()[:n].reshape(theta.shape)
               for f in (ids.TH, ids.QV, ids.QC, ids.QR)]
        return dict(dth=out[0], qv=out[1], qc=out[2], qr=out[3])

    def saturation(self, T, p):
        """qvs at temperature T and pressure p: the vapour above which the scheme condenses (qc = qr = 0,
        so only the adjustment can act), found by bisection."""
        T, p = np.broadcast_arrays(np.asarray(T, float), np.asarray(p, float))
        exner = (p / P0) ** (RD / CP_REF)
        lo, hi = np.zeros(T.shape), np.full(T.shape, 0.2)
        for _ in range(64):
            mid = 0.5 * (lo + hi)
            cond = self.run(T / exner, exner, mid, 0.0, 0.0)["qc"] > 0
            hi = np.where(cond, mid, hi)
            lo = np.where(cond, lo, mid)
        return hi, exner


@pytest.fixture(params=BACKENDS)
def kind(request):
    return request.param


def test_saturation_vapour_pressure_matches_iapws(kind):
    g = KesslerGrid(kind)
    tc = np.array(sorted(IAPWS_ES))
    p = 1.0e5
    qs, _ = g.saturation(tc + 273.15, p)
    es = qs * p / (EPS + qs)   # invert qvs = eps e_s / (p - e_s)
    ref = np.array([IAPWS_ES[t] for t in tc])
    rel = es / ref - 1.0
    print(f"{kind}: e_s / IAPWS - 1 at {tc} C: {np.array2string(rel, precision=5)}")
    assert np.all(np.abs(rel) < 2e-3)   # Bolton's stated accuracy is 0.1 %; 1 % in any constant fails


def test_condensation_is_a_newton_step_of_saturation(kind):
    g = KesslerGrid(kind)
    T = np.array([280.0, 295.0, 303.0, 285.0])
    p = np.array([9.0e4, 9.5e4, 1.0e5, 7.0e4])
    qs, exner = g.saturation(T, p)
    dT = 0.01
    qs_hi, _ = g.saturation(T + dT, p)
    qs_lo, _ = g.saturation(T - dT, p)
    dqs_dT = (qs_hi - qs_lo) / (2 * dT)
    qv = 1.02 * qs
    out = g.run(T / exner, exner, qv, 0.0, 0.0)
    C = out["qc"]
    assert np.all(C > 0)
    np.testing.assert_allclose(out["qv"], qv - C, rtol=1e-14)
    # latent heating: d theta = Lv / (cp pi) C
    lv_cp = out["dth"] * exner / C
    assert np.all(np.abs(lv_cp / (LV_PUB / CP_PUB) - 1) < 2e-3)
    # one Newton step: C (1 + (Lv/cp) dqs/dT) = qv - qs
    phi = (qv - qs) / C - 1.0
    np.testing.assert_allclose(phi, lv_cp * dqs_dT, rtol=1e-6)
    # subsaturated with cloud: the same step evaporates cloud, never more than there is
    out = g.run(T / exner, exner, 0.98 * qs, np.array([1e-5, 1e-4, 5e-4, 1e-6]), 0.0)
    C = out["qv"] - 0.98 * qs
    np.testing.assert_allclose(C, np.minimum([1e-5, 1e-4, 5e-4, 1e-6], 0.02 * qs / (1 + phi)), rtol=1e-9)


def test_autoconversion_rate(kind):
    g = KesslerGrid(kind)
    T, p = 290.0, 9.0e4
    qs, exner = g.saturation(T, p)
    qc = np.array([0.5e-3, 0.999e-3, 1.001e-3, 1.5e-3, 3e-3, 1e-2])
    out = g.run(T / exner, exner, 0.5 * qs, qc, 0.0)  # no rain: accretion and evaporation are off
    expect = np.minimum(K1 * np.maximum(qc - A_AUTO, 0.0) * DT, qc)
    np.testing.assert_allclose(out["qr"], expect, rtol=1e-12, atol=1e-20)


@pytest.mark.parametrize("rho", [1.1, 0.55])
def test_accretion_rate(kind, rho):
    g = KesslerGrid(kind, rho=rho)
    T, p = 290.0, 9.0e4
    qs, exner = g.saturation(T, p)
    qr = np.array([1e-11, 1e-6, 1e-4, 1e-3, 4e-3, 1e-2])[:, None]
    qc = np.array([2e-4, 8e-4, 1e-3])[None, :]             # at or below the autoconversion threshold
    out = g.run(T / exner, exner, 1.05 * qs, qc, qr)     # supersaturated: evaporation clamps to zero
    rate = K2 * qc * qr ** ACC_EXP                         # Klemp & Wilhelmson, s^-1
    expect = qr + np.minimum(rho * rate * DT, qc)          # the reference's extra rhou[k] factor
    np.testing.assert_allclose(out["qr"], expect, rtol=1e-12)


def _evap_published(qr, qv, qs, rho, p):
    rq = G_PER_CM3 * rho * qr                              # rho qr in g cm^-3
    vent = VENT0 + VENT1 * rq ** VENT_EXP
    e = (1.0 - qv / qs) * vent * rq ** EVAP_EXP / (EVAP_D1 + EVAP_D2 / (MB_PER_PA * p * qs)) / (G_PER_CM3 * rho)
    return np.minimum(np.maximum(e * DT, 0.0), qr)


EVAP_REGIMES = [(295.0, 9.5e4), (250.0, 4.0e4)]
EVAP_QR = np.array([1e-8, 1e-5, 1e-3, 1e-2])
EVAP_RH = np.array([0.3, 0.8])


@pytest.mark.parametrize("rho", [1.1, 0.5])
@pytest.mark.parametrize("regime", EVAP_REGIMES)
def test_rain_evaporation_rate(kind, rho, regime):
    T, p = regime
    g = KesslerGrid(kind, rho=rho)
    qs, exner = g.saturation(T, p)
    qr = EVAP_QR[:, None]
    qv = EVAP_RH[None, :] * qs
    out = g.run(T / exner, exner, qv, 0.0, qr)
    E = qr - out["qr"]
    assert np.all(E > 0)
    np.testing.assert_allclose(out["qv"] - qv, E, rtol=0, atol=4e-16 * qv.max())   # water is conserved
    expect = _evap_published(qr, qv, qs, rho, p)
    rel = E / expect - 1
    print(f"{kind} T={T} p={p} rho={rho}: E/E_KW78 - 1 in [{rel.min():+.2e}, {rel.max():+.2e}]")
    assert np.all(np.abs(rel) < 1e-3)
    # evaporative cooling
    assert np.all(np.abs(-out["dth"] * exner / E / (LV_PUB / CP_PUB) - 1) < 2e-3)


def test_evaporation_cases_resolve_every_constant():
    """A 1 % change in any constant of the evaporation moves some case of test_rain_evaporation_rate by
    more than 2.5x its 0.1 % tolerance (computed on the published expression itself, with a typical
    saturation mixing ratio per regime)."""
    global VENT0, VENT1, VENT_EXP, EVAP_EXP, EVAP_D1, EVAP_D2
    names = ["VENT0", "VENT1", "VENT_EXP", "EVAP_EXP", "EVAP_D1", "EVAP_D2"]
    regimes = [(295.0, 9.5e4, 0.0175), (250.0, 4.0e4, 0.0015)]
    base = {n: globals()[n] for n in names}

    def all_cases():
        out = []
        for T, p, qs in regimes:
            for rho in (1.1, 0.5):
                out.append(_evap_published(EVAP_QR[:, None], EVAP_RH[None, :] * qs, qs, rho, p).ravel())
        return np.concatenate(out)

    ref = all_cases()
    try:
        for n in names:
            globals()[n] = base[n] * 1.01
            moved = np.abs(all_cases() / ref - 1).max()
            globals()[n] = base[n]
            assert moved > 2.5e-3, (n, moved)
    finally:
        for n in names:
            globals()[n] = base[n]


def test_si_constants_of_the_reference_follow_from_the_cgs_forms():
    """kessler.cpp:70-76 writes the evaporation in SI; the constants are the cgs ones converted."""
    # ventilation: 124.9 (1e-3 rho qr)^0.2046 = 30.3922 (rho qr)^0.2046
    assert abs(VENT1 * G_PER_CM3 ** VENT_EXP / 30.3922 - 1) < 2e-5
    # (1/rho_c)(rho_c qr)^0.525 = 10^1.425 (1/rho)(rho qr)^0.525, so dividing the denominator by 10^1.425
    # and converting mb to Pa gives the reference's 2.03e4 and 9.58e6 (three significant figures)
    f = G_PER_CM3 ** (EVAP_EXP - 1.0)
    assert round(EVAP_D1 / f, -2) == 2.03e4
    assert round(EVAP_D2 / MB_PER_PA / f, -4) == 9.58e6
    # fall speed: 0.001 turns rho qr from kg m^-3 into g cm^-3 (kessler.cpp:237)
    assert G_PER_CM3 == 0.001


@pytest.mark.parametrize("backend", BACKENDS)
def test_fall_speed_published_form(backend):
    """calculate_rain_fall_speed_kessler (kessler.cpp:222-252), through advect_microphysics_cell, which
    stores it (advection.cpp:456-457): V = 36.34 (rho_c qr)^0.1364 (rho_1/rho_k)^(1/2), zero at or below
    qr = 1e-12, level 0 a copy of level 1."""
    c = zero_case(nx=16, ny=16, nz=12, microphysics=1, dt=DT)
    rho = np.linspace(1.15, 0.45, c.nz)
    c.columns["rhou"] = rho
    rng = np.random.Generator(np.random.PCG64(11))
    qr = 10.0 ** rng.uniform(-13, -1.5, size=(c.nx, c.ny, c.nz))
    qr[:, :, 5] = 1e-12
    c.fields["qr"] = qr
    b = make_backend(backend, c)
    b.advect_moisture(1.0 / 3.0, 3, c.nx - 3, 3, c.ny - 3)
    if backend == "hip":
        b.sync()
    vt = b.get_field(ids.VT, 0)[3:-3, 3:-3]
    q = qr[3:-3, 3:-3]
    k = np.arange(c.nz)
    expect = np.where(q > 1e-12, VT_C * (G_PER_CM3 * rho[None, None, :] * q) ** VT_EXP * np.sqrt(rho[1] / rho)[None, None, :], 0.0)
    np.testing.assert_allclose(vt[:, :, 1:-1], expect[:, :, 1:-1], rtol=1e-13, atol=0)
    np.testing.assert_array_equal(vt[:, :, 0], vt[:, :, 1])
    # 1 g m^-3 of rain falls at about 5.5 m/s at the lowest level
    assert abs(VT_C * (1e-6) ** VT_EXP - 5.52) < 0.01
