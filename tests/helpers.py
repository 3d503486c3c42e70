"""Shared test helpers: state construction and backend selection (oracle on the CPU, HIP on the GPU)."""
import numpy as np
import pytest

from model_amd import configs, ids


def zero_case(nx=16, ny=16, nz=12, microphysics=0, unit_density=False, dx=1000.0, dz=250.0, dt=2.0, f0=0.0,
              rayleigh_level=None):
    """A small case with every 3-D field zero (basic state included) and pbar = pib."""
    c = configs.small_case(nx=nx, ny=ny, nz=nz, microphysics=microphysics, dx=dx, dz=dz, dt=dt, f0=f0,
                           rayleigh_level=rayleigh_level)
    for k in list(c.fields):
        c.fields[k] = np.zeros_like(c.fields[k])
    c.fields["pbar"][:] = c.columns["pib"][None, None, :]
    c.columns["f"] = np.full(c.ny, f0)
    if unit_density:
        for n in ("rhou", "rhow", "one_d_rhou", "one_d_rhow"):
            c.columns[n] = np.ones(c.nz)
    return c


def index_grids(case):
    i = np.arange(case.nx, dtype=np.float64)[:, None, None]
    j = np.arange(case.ny, dtype=np.float64)[None, :, None]
    k = np.arange(case.nz, dtype=np.float64)[None, None, :]
    return (np.broadcast_to(i, (case.nx, case.ny, case.nz)), np.broadcast_to(j, (case.nx, case.ny, case.nz)),
            np.broadcast_to(k, (case.nx, case.ny, case.nz)))


def make_backend(kind, case):
    if kind == "oracle":
        from oracle.pyoracle import Oracle
        return Oracle(case)
    from model_amd.core import DynCore
    return DynCore(case)


BACKENDS = ["oracle", pytest.param("hip", marks=pytest.mark.gpu)]


def rel_err(a, b):
    s = np.abs(b).max()
    d = np.abs(a - b).max()
    return d / s if s > 0 else d
