"""Synthetic configurations of the anelastic dynamical core.

The reference initialises from NetCDF files or code tied to latitude; these builders create the same
kind of input directly in memory: the 1-D base state (initializer.cpp:840-928, 1340-1380), the 3-D
basic state (ubar ... pbar) and the initial perturbation, on the reference's grid conventions:

  * NX = nx + 6 (three periodic halo columns on each side, boundaries.cpp:562-574),
  * NZ counts the ghost levels k = 0 and k = NZ-1 (boundaries.cpp:304-326),
  * layout [i][j][k], k fastest (model.h:50).

Noise is drawn once on the host in FP64 from numpy's PCG64 with seed 20261202 + case id, so every
path receives identical bytes.  All three rotating buffers (M, current, forecast) start equal.
"""
from __future__ import annotations

import dataclasses
import math

import numpy as np

from . import ids

# physical constants, model.h:250-259
RD, CP, CV, P0, GRAV, LV = 287.0, 1004.0, 717.0, 100000.0, 9.8, 2.5e6
PRESSFC = 100200.0
SEED_BASE = 20261202


@dataclasses.dataclass
class Case:
    name: str
    nx: int            # NX including the 6 halo columns
    ny: int
    nz: int            # NZ including the 2 ghost levels
    dx: float
    dz: float
    dt: float
    microphysics: int  # 0 dry, 1 Kessler
    rayleigh_level: int
    steps: int
    columns: dict      # column name -> array (nz, or ny for "f")
    fields: dict       # field name -> [NX, NY, NZ] float64 (initial state of every buffer)
    sponge: int = 5
    notes: str = ""

    @property
    def interior_points(self) -> int:
        """Gridpoints per step as the throughput metric counts them: nx_interior * ny * nz."""
        return (self.nx - 6) * self.ny * self.nz

    def grid_tuple(self):
        return (self.nx, self.ny, self.nz, self.dx, self.dz, self.dt, self.microphysics,
                self.rayleigh_level, self.sponge, self.sponge)


# --------------------------------------------------------------------------------------------------
# 1-D base state
# --------------------------------------------------------------------------------------------------
def vertical_levels(nz: int, dz: float, height_lowest_level: float | None = None):
    """setup_vertical_height_levels + stretched_grid (initializer.cpp:840-850, 1340-1380)."""
    z0 = dz if height_lowest_level is None else height_lowest_level
    k = np.arange(nz, dtype=np.float64)
    zu = (k - 0.5) * dz
    zw = (k - 1.0) * dz
    nt, lev = nz - 1, 1
    c2 = (1 - z0 / (dz * lev)) / ((nt - 1) * dz - dz * lev)
    c1 = z0 / (dz * lev) - c2 * dz * lev
    full = (k - 1.0) * dz
    half = (k - 0.5) * dz
    zsw = (c1 + c2 * full) * full
    mw = 1 / (c1 + 2 * c2 * full)
    zsu = (c1 + c2 * half) * half
    mu = 1 / (c1 + 2 * c2 * half)
    return dict(zu=zu, zw=zw, zsu=zsu, zsw=zsw, mu=mu, mw=mw)


def base_state(tb: np.ndarray, qb: np.ndarray, zsu: np.ndarray):
    """initialize_vertical_basic_state2 (initializer.cpp:858-928), stretched branch."""
    nz = tb.size
    tb = tb.astype(np.float64).copy()
    qb = qb.astype(np.float64).copy()
    tb[0], tb[nz - 1] = tb[1], tb[nz - 2]
    qb[0], qb[nz - 1] = qb[1], qb[nz - 2]
    tbw = np.empty(nz)
    tbw[0] = 0.5 * (tb[0] + tb[1])
    tbw[1:] = 0.5 * (tb[1:] + tb[:-1])
    tbv = tb * (1.0 + 0.61 * qb)
    pisfc = (PRESSFC / P0) ** (RD / CP)
    pib = np.empty(nz)
    pib[1] = pisfc - GRAV * (zsu[1] - 0) / (CP * tbv[1])
    pib[0] = pisfc + GRAV * (0 - zsu[0]) / (CP * tbv[0])
    for k in range(2, nz):
        pib[k] = pib[k - 1] - GRAV * (zsu[k] - zsu[k - 1]) / (CP * (0.5 * (tbv[k] + tbv[k - 1])))
    rhou = P0 * pib ** (CV / RD) / (RD * tbv)
    rhow = np.empty(nz)
    rhow[1:] = 0.5 * (rhou[1:] + rhou[:-1])
    tbv[0], tbv[nz - 1] = tbv[1], tbv[nz - 2]
    tbw[0], tbw[nz - 1] = tbw[1], tbw[nz - 2]
    pib[0], pib[nz - 1] = pib[1], pib[nz - 2]
    rhow[0] = rhow[1]
    return dict(tb=tb, qb=qb, tbw=tbw, tbv=tbv, pib=pib, rhou=rhou, rhow=rhow,
                one_d_rhow=1.0 / rhow, one_d_rhou=1.0 / rhou)


def qv_sat(temperature, pressure):
    """get_QV_Sat (initializer.cpp:1387-1393)."""
    esl = 611.2 * np.exp(17.67 * (temperature - 273.15) / (temperature - 29.65))
    return 0.62197 * esl / (pressure - esl)


def _columns(nz, ny, dz, tb, qb, f=0.0):
    lv = vertical_levels(nz, dz)
    bs = base_state(tb, qb, lv["zsu"])
    cols = {**lv, **bs}
    cols["f"] = np.full(ny, float(f))
    return {n: np.ascontiguousarray(cols[n], dtype=np.float64) for n in ids.COLUMN_NAMES}


def _empty_fields(NX, NY, NZ):
    return {n: np.zeros((NX, NY, NZ)) for n in ids.FIELD_NAMES if n != "accrain"}


def _finish(fields, NX, NY, NZ):
    """Periodic halo columns (boundaries.cpp:562-574) and ghost levels k=0, NZ-1 for the initial state."""
    for name, a in fields.items():
        if a.ndim != 3:
            continue
        a[0:3] = a[NX - 6:NX - 3]
        a[NX - 3:NX] = a[3:6]
        if name in ("w", "wbar"):
            a[:, :, 0] = 0.0
            a[:, :, NZ - 1] = a[:, :, NZ - 2] if name == "wbar" else 0.0
        else:
            a[:, :, 0] = a[:, :, 1]
            a[:, :, NZ - 1] = a[:, :, NZ - 2]
    fields["accrain"] = np.zeros((NX, NY))
    return fields


def _coords(NX, NY, cols, dx):
    x = (np.arange(NX) - 3.0) * dx
    y = np.arange(NY) * dx
    return x, y, cols["zsu"]


# --------------------------------------------------------------------------------------------------
# C1: plain FP64 dry warm bubble, 64x64x40
# --------------------------------------------------------------------------------------------------
def warm_bubble(nx=64, ny=64, nz=40, dx=1000.0, dz=250.0, dt=2.0, steps=100, seed_offset=1) -> Case:
    NX, NY, NZ = nx + 6, ny, nz
    z = vertical_levels(NZ, dz)["zsu"]
    n2 = 0.01 ** 2
    tb = 300.0 * np.exp(n2 * z / GRAV)
    cols = _columns(NZ, NY, dz, tb, np.zeros(NZ), f=0.0)
    fields = _empty_fields(NX, NY, NZ)
    fields["pbar"][:] = cols["pib"][None, None, :]
    x, y, zs = _coords(NX, NY, cols, dx)
    xc, yc, zc = 0.5 * (nx - 1) * dx, 0.5 * (ny - 1) * dx, 1500.0
    X, Y, Z = np.meshgrid(x, y, zs, indexing="ij")
    r = np.sqrt((X - xc) ** 2 + (Y - yc) ** 2 + (Z - zc) ** 2) / 2000.0
    th = np.where(r <= 1.0, 2.0 * np.cos(0.5 * np.pi * r) ** 2, 0.0)
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + seed_offset))
    noise = rng.uniform(-0.01, 0.01, size=(nx, NY, NZ - 2))
    th[3:NX - 3, :, 1:NZ - 1] += noise
    fields["th"] = th
    return Case("warm_bubble", NX, NY, NZ, dx, dz, dt, 0, NZ, steps, cols, _finish(fields, NX, NY, NZ),
                notes="2 K cos^2 bubble r=2 km at z=1.5 km + U(-0.01,0.01) K noise, N=0.01, no wind, f=0")


# --------------------------------------------------------------------------------------------------
# C2: moist Kessler supercell-style run, 256x256x64
# --------------------------------------------------------------------------------------------------
def supercell(nx=256, ny=256, nz=64, dx=1000.0, dz=250.0, dt=2.0, steps=1800, seed_offset=2,
              bubble=True, noise=0.0) -> Case:
    NX, NY, NZ = nx + 6, ny, nz
    z = vertical_levels(NZ, dz)["zsu"]
    ztr, ttr, temptr = 12000.0, 343.0, 213.0
    zc = np.clip(z, 0.0, None)
    tb = np.where(z < ztr, 300.0 + (ttr - 300.0) * (zc / ztr) ** 1.25,
                  ttr * np.exp(GRAV * (z - ztr) / (CP * temptr)))
    rh = np.where(z < ztr, 1.0 - 0.75 * (zc / ztr) ** 1.25, 0.25)
    qb = np.zeros(NZ)
    # fixed-point iteration between qv (needs T, p) and the virtual-temperature pressure profile
    for _ in range(4):
        bs = base_state(tb, qb, z)
        temp = tb * bs["pib"]
        pres = P0 * bs["pib"] ** (CP / RD)
        qb = np.minimum(rh * qv_sat(temp, pres), 0.014)
    cols = _columns(NZ, NY, dz, tb, qb, f=0.0)
    fields = _empty_fields(NX, NY, NZ)
    fields["pbar"][:] = cols["pib"][None, None, :]
    fields["ubar"][:] = (20.0 * np.tanh(z / 3000.0))[None, None, :]
    x, y, zs = _coords(NX, NY, cols, dx)
    X, Y, Z = np.meshgrid(x, y, zs, indexing="ij")
    if bubble:
        xc, yc = 0.5 * (nx - 1) * dx, 0.5 * (ny - 1) * dx
        r = np.sqrt(((X - xc) / 10000.0) ** 2 + ((Y - yc) / 10000.0) ** 2 + ((Z - 1400.0) / 1400.0) ** 2)
        fields["th"] = np.where(r <= 1.0, 3.0 * np.cos(0.5 * np.pi * r) ** 2, 0.0)
    if noise > 0:
        rng = np.random.Generator(np.random.PCG64(SEED_BASE + seed_offset))
        fields["th"][3:NX - 3, :, 1:NZ - 1] += rng.uniform(-noise, noise, size=(nx, NY, NZ - 2))
    ray = int(np.searchsorted(cols["zu"], 12000.0))
    return Case("supercell", NX, NY, NZ, dx, dz, dt, 1, ray, steps, cols, _finish(fields, NX, NY, NZ),
                notes="theta 300->343 K at 12 km (z^1.25), isothermal above; RH=1-0.75(z/12km)^1.25, qv<=14 g/kg "
                      "(4 fixed-point sweeps); u=20 tanh(z/3 km); 3 K ellipsoid 10 km x 1.4 km at z=1.4 km; "
                      "Rayleigh damping from the first level with zu>=12 km")


# --------------------------------------------------------------------------------------------------
# C3/C4: dry vortex in shear on an f-plane
# --------------------------------------------------------------------------------------------------
def vortex_in_shear(nx=192, ny=192, nz=40, dx=20000.0, dz=500.0, dt=60.0, steps=360, seed_offset=4,
                    f0=4.4e-5, zeta0=1e-4, radius=300e3, hz=5000.0, noise_rms=0.1) -> Case:
    NX, NY, NZ = nx + 6, ny, nz
    z = vertical_levels(NZ, dz)["zsu"]
    n2 = 0.01 ** 2
    tb = 300.0 * np.exp(n2 * np.clip(z, 0.0, None) / GRAV)
    cols = _columns(NZ, NY, dz, tb, np.zeros(NZ), f=f0)
    fields = _empty_fields(NX, NY, NZ)
    fields["pbar"][:] = cols["pib"][None, None, :]
    x, y, zs = _coords(NX, NY, cols, dx)
    # basic-state u(y,z): a horizontal shear strip plus a piecewise-linear vertical profile with shears
    # -1e-3 below and -2e-3 above 5 km.  The strip is u = -zeta0 W tanh((y - yc)/W): its vorticity
    # -du/dy peaks at zeta0 on the centre line; W = 150 km gives the width of the reference's 18-21 N
    # strip (data_initializer.cpp:350, about 330 km) and a jet of +-15 m/s.
    yc = 0.5 * (ny - 1) * dx
    width = 150e3
    uy = -zeta0 * width * np.tanh((y - yc) / width)
    uz = np.where(zs < 5000.0, -1e-3 * zs, -1e-3 * 5000.0 - 2e-3 * (zs - 5000.0))
    fields["ubar"][:] = (uy[:, None] + uz[None, :])[None, :, :]
    # axisymmetric Gaussian-vorticity vortex: v_t(r) = zeta0 R^2/(2r) (1 - exp(-r^2/R^2)) e^{-z/H}
    xc = 0.5 * (nx - 1) * dx
    X, Y, Z = np.meshgrid(x, y, zs, indexing="ij")
    Xu, Yv = X - 0.5 * dx, Y - 0.5 * dx  # u lives on the west face, v on the south face
    decay = np.exp(-np.clip(Z, 0.0, None) / hz)

    def vt(r):
        r = np.maximum(r, 1.0)
        return zeta0 * radius ** 2 / (2 * r) * (1 - np.exp(-(r / radius) ** 2))

    ru = np.hypot(Xu - xc, Y - yc)
    rv = np.hypot(X - xc, Yv - yc)
    fields["u"] = -vt(ru) * (Y - yc) / np.maximum(ru, 1.0) * decay
    fields["v"] = vt(rv) * (X - xc) / np.maximum(rv, 1.0) * decay
    # warm core in thermal-wind balance with the gradient wind, to leading order:
    # d theta'/dr = (theta0/g) (f + 2 v/r) d v/dz  with d v/dz = -v/H  (integrated numerically)
    rr = np.linspace(0, 3 * radius * 2, 4000)
    vr = vt(rr)
    dthdr = (300.0 / GRAV) * (f0 + 2 * vr / np.maximum(rr, 1.0)) * (-vr / hz)
    th_r = np.concatenate([[0.0], np.cumsum(0.5 * (dthdr[1:] + dthdr[:-1]) * np.diff(rr))])
    th_r -= th_r[-1]
    rc = np.hypot(X - xc, Y - yc)
    fields["th"] = np.interp(rc, rr, th_r) * decay
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + seed_offset))
    fields["th"][3:NX - 3, :, 1:NZ - 1] += rng.normal(0.0, noise_rms, size=(nx, NY, NZ - 2))
    return Case("vortex_in_shear", NX, NY, NZ, dx, dz, dt, 0, NZ, steps, cols, _finish(fields, NX, NY, NZ),
                notes="f-plane f=4.4e-5; Gaussian-vorticity vortex peak 1e-4 s^-1, R=300 km, e-folding 5 km; "
                      "thermal-wind warm core; u(y,z) shear strip + vertical shears -1e-3/-2e-3 s^-1 below/above 5 km; "
                      "0.1 K RMS noise")


# --------------------------------------------------------------------------------------------------
# C5: large moist tropical vortex, 512x512x80
# --------------------------------------------------------------------------------------------------
def tropical_vortex(nx=512, ny=512, nz=80, dx=2000.0, dz=250.0, dt=4.0, steps=600, seed_offset=5) -> Case:
    NX, NY, NZ = nx + 6, ny, nz
    z = vertical_levels(NZ, dz)["zsu"]
    zc = np.clip(z, 0.0, None)
    tb = 300.0 * np.exp(0.011 ** 2 * zc / GRAV)
    qb = 0.018 * np.exp(-zc / 2500.0)
    cols = _columns(NZ, NY, dz, tb, qb, f=0.0)
    fields = _empty_fields(NX, NY, NZ)
    fields["pbar"][:] = cols["pib"][None, None, :]
    x, y, zs = _coords(NX, NY, cols, dx)
    xc, yc = 0.5 * (nx - 1) * dx, 0.5 * (ny - 1) * dx
    X, Y, Z = np.meshgrid(x, y, zs, indexing="ij")
    rm, vmax, hz = 50e3, 15.0, 8000.0
    decay = np.exp(-np.clip(Z, 0.0, None) / hz)

    def vt(r):
        return vmax * (r / rm) * np.exp(1.0 - r / rm)

    ru = np.hypot(X - 0.5 * dx - xc, Y - yc)
    rv = np.hypot(X - xc, Y - 0.5 * dx - yc)
    fields["u"] = -vt(ru) * (Y - yc) / np.maximum(ru, 1.0) * decay
    fields["v"] = vt(rv) * (X - 0.5 * dx - xc + 0.5 * dx) / np.maximum(rv, 1.0) * decay
    # cyclostrophic warm core: d theta'/dr = (theta0/g) (v^2/r)/H, integrated outward, zero far away
    rr = np.linspace(0, 12 * rm, 6000)
    vr = vt(rr)
    dthdr = -(300.0 / GRAV) * (vr ** 2 / np.maximum(rr, 1.0)) / hz
    th_r = np.concatenate([[0.0], np.cumsum(0.5 * (dthdr[1:] + dthdr[:-1]) * np.diff(rr))])
    th_r -= th_r[-1]
    fields["th"] = np.interp(np.hypot(X - xc, Y - yc), rr, th_r) * decay
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + seed_offset))
    low = (zs < 1000.0) & (zs > 0)
    kk = np.nonzero(low)[0]
    fields["th"][3:NX - 3, :, kk] += rng.uniform(-0.2, 0.2, size=(nx, NY, kk.size))
    ray = int(np.searchsorted(cols["zu"], 16000.0))
    return Case("tropical_vortex", NX, NY, NZ, dx, dz, dt, 1, ray, steps, cols, _finish(fields, NX, NY, NZ),
                notes="theta 300 K, N=0.011; qv=18 g/kg exp(-z/2.5 km); no mean wind; 15 m/s vortex at 50 km, "
                      "cyclostrophic warm core, 0.2 K noise below 1 km")


def small_case(nx=16, ny=16, nz=12, microphysics=0, seed=7, amplitude=1.0, dx=1000.0, dz=250.0, dt=2.0,
               rayleigh_level=None, f0=1e-4) -> Case:
    """Random, well-scaled state on a tiny grid for operator-level parity tests (all fields non-trivial)."""
    NX, NY, NZ = nx + 6, ny, nz
    rng = np.random.Generator(np.random.PCG64(SEED_BASE + 100 + seed))
    z = vertical_levels(NZ, dz)["zsu"]
    tb = 300.0 * np.exp(1e-4 * np.clip(z, 0, None) / GRAV)
    qb = (0.014 * np.exp(-np.clip(z, 0, None) / 2500.0)) if microphysics else np.zeros(NZ)
    cols = _columns(NZ, NY, dz, tb, qb, f=f0)
    cols["f"] = f0 * (1.0 + 0.1 * rng.standard_normal(NY))
    fields = _empty_fields(NX, NY, NZ)
    shape = (NX, NY, NZ)
    a = amplitude
    fields["u"] = a * 5.0 * rng.standard_normal(shape)
    fields["v"] = a * 5.0 * rng.standard_normal(shape)
    fields["w"] = a * 2.0 * rng.standard_normal(shape)
    fields["th"] = a * 1.0 * rng.standard_normal(shape)
    fields["ubar"] = 10.0 + 3.0 * rng.standard_normal(shape)
    fields["vbar"] = -2.0 + 3.0 * rng.standard_normal(shape)
    fields["wbar"] = 0.1 * rng.standard_normal(shape)
    fields["thbar"] = 0.5 * rng.standard_normal(shape)
    fields["pbar"] = cols["pib"][None, None, :] + 1e-4 * rng.standard_normal(shape)
    if microphysics:
        fields["qv"] = 1e-3 * rng.standard_normal(shape)
        fields["qc"] = np.abs(1.5e-3 * rng.standard_normal(shape))
        fields["qr"] = np.abs(1.5e-3 * rng.standard_normal(shape))
        fields["qbar"] = 5e-4 * rng.standard_normal(shape)
        fields["pi"] = 50.0 * rng.standard_normal(shape)
    ray = NZ if rayleigh_level is None else rayleigh_level
    fields["accrain"] = np.zeros((NX, NY))
    return Case(f"small{nx}x{ny}x{nz}{'m' if microphysics else 'd'}", NX, NY, NZ, dx, dz, dt, microphysics, ray, 1,
                cols, fields, notes="random operator-test state")


BUILDERS = {"warm_bubble": warm_bubble, "supercell": supercell, "vortex_in_shear": vortex_in_shear,
            "tropical_vortex": tropical_vortex}
