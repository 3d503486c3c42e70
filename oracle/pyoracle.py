"""ctypes wrapper of the oracle (oracle/dyncore_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU-baseline
leg, as the checker or the timed CPU restatement; never by the model_amd package.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}

_dp = ctypes.POINTER(ctypes.c_double)


class OrcGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
                ("microphysics", ctypes.c_int), ("rayleigh_level", ctypes.c_int),
                ("sponge_north", ctypes.c_int), ("sponge_south", ctypes.c_int)]


def build():
    """Compile the oracle with its Makefile (parity and timed variants)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE, "-B"], check=True)


def lib(variant: str = "parity"):
    name = {"parity": "liboracle.so", "omp": "liboracle_omp.so", "tlm": "liboracle_tlm.so", "adj": "liboracle_adj.so"}[variant]
    if variant in _LIBS:
        return _LIBS[variant]
    path = os.path.join(_HERE, "_build", name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built (run `make -C oracle`)")
    L = ctypes.CDLL(path)
    if variant == "adj":
        L.orc_adj_create.restype = ctypes.c_void_p
        L.orc_adj_create.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_adj_destroy.argtypes = [ctypes.c_void_p]
        L.orc_adj_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_adj_error.argtypes = [ctypes.c_void_p]
        L.orc_adj_set_column.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_adj_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_adj_set_initial.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        L.orc_adj_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
        L.orc_adj_gradient.argtypes = [ctypes.c_void_p] + [_dp] * 8
        L.orc_adj_tape_size.restype = ctypes.c_longlong
        _LIBS[variant] = L
        return L
    if variant == "tlm":
        L.orc_tlm_create.restype = ctypes.c_void_p
        L.orc_tlm_create.argtypes = [ctypes.POINTER(OrcGrid)]
        L.orc_tlm_destroy.argtypes = [ctypes.c_void_p]
        L.orc_tlm_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        L.orc_tlm_error.argtypes = [ctypes.c_void_p]
        L.orc_tlm_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_tlm_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_tlm_set_column.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
        _LIBS[variant] = L
        return L
    L.orc_create.restype = ctypes.c_void_p
    L.orc_create.argtypes = [ctypes.POINTER(OrcGrid)]
    L.orc_destroy.argtypes = [ctypes.c_void_p]
    L.orc_error.argtypes = [ctypes.c_void_p]
    for fn in ("orc_set_column", "orc_get_column"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    for fn in ("orc_set_field", "orc_get_field"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp]
    L.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    for fn in ("orc_advect_uvw", "orc_advect_theta", "orc_advect_moisture", "orc_pgf"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4
    for fn in ("orc_zero_moisture", "orc_kessler"):
        getattr(L, fn).argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
    L.orc_rayleigh.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
    L.orc_solve_pressure.argtypes = [ctypes.c_void_p, ctypes.c_double]
    for fn in ("orc_apply_bc", "orc_apply_bc_micro", "orc_advance_inner", "orc_advance_outer",
               "orc_micro_advance_inner", "orc_micro_advance_outer"):
        getattr(L, fn).argtypes = [ctypes.c_void_p]
    L.orc_fft1d.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
    L.orc_fft2d.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int]
    for fn in ("orc_interp5", "orc_interp3"):
        getattr(L, fn).argtypes = [_dp, ctypes.c_double]
        getattr(L, fn).restype = ctypes.c_double
    L.orc_interp2.argtypes = [_dp]
    L.orc_interp2.restype = ctypes.c_double
    L.orc_poisson_rhs.argtypes = [ctypes.c_void_p, _dp, _dp]
    L.orc_kessler_point.argtypes = [ctypes.c_double] * 9 + [_dp] * 4
    L.orc_kessler_point.restype = ctypes.c_int
    L.orc_fall_speed.argtypes = [ctypes.c_double] * 3
    L.orc_fall_speed.restype = ctypes.c_double
    L.orc_set_turbulence.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc_calculate_diff_tend.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
    for fn in ("orc_apply_velocity_diffusion", "orc_apply_scalar_diffusion", "orc_apply_moisture_diffusion"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4
    L.orc_get_turbulence.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp]
    L.orc_set_saturated.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
    _LIBS[variant] = L
    return L


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


class Oracle:
    """One oracle state; mirrors the reference's serial globals."""

    def __init__(self, case, variant: str = "parity"):
        self.L = lib(variant)
        self.case = case
        g = OrcGrid(case.nx, case.ny, case.nz, case.dx, case.dz, case.dt, case.microphysics,
                    case.rayleigh_level, case.sponge, case.sponge)
        self.h = self.L.orc_create(ctypes.byref(g))
        from model_amd import ids  # identifiers only
        self.ids = ids
        for name, arr in case.columns.items():
            self.set_column(ids.COLUMN_ID[name], arr)
        for name, arr in case.fields.items():
            f = ids.FIELD_ID[name]
            if f in ids.PROGNOSTIC:
                for role in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
                    self.set_field(f, role, arr)
            else:
                self.set_field(f, 0, arr)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_column(self, c, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.L.orc_set_column(self.h, c, _ptr(v))

    def get_column(self, c):
        n = self.case.ny if c == self.ids.COL_F else self.case.nz
        out = np.empty(n)
        self.L.orc_get_column(self.h, c, _ptr(out))
        return out

    def set_field(self, f, role, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        self.L.orc_set_field(self.h, f, role, _ptr(v))

    def get_field(self, f, role=1):
        c = self.case
        shape = (c.nx, c.ny) if f == self.ids.ACCRAIN else (c.nx, c.ny, c.nz)
        out = np.empty(shape)
        self.L.orc_get_field(self.h, f, role, _ptr(out))
        return out

    def step(self, n=1):
        self.L.orc_step(self.h, n)

    def set_turbulence(self, on=True, landsea=None):
        ls = None if landsea is None else np.ascontiguousarray(landsea, dtype=np.float64)
        self._landsea = ls
        self.L.orc_set_turbulence(self.h, int(on), None if ls is None else _ptr(ls))

    def set_saturated(self, flags):
        """Kessler's saturation flags (nx, ny, nz; nonzero = saturated), as the next calculate_diff_tend
        reads them; turbulence must be on."""
        f = np.ascontiguousarray(flags, dtype=np.uint8)
        assert f.shape == (self.case.nx, self.case.ny, self.case.nz)
        self.L.orc_set_saturated(self.h, f.ctypes.data)

    def get_turbulence(self, which):
        """0 Kmix, 1 KHmix, 2-4 u/v/w friction, 5-8 theta/qv/qc/qr diffusion tendencies, 9 saturation flag."""
        c = self.case
        out = np.zeros((c.nx, c.ny, c.nz))
        self.L.orc_get_turbulence(self.h, which, _ptr(out))
        return out

    @property
    def error(self):
        return self.L.orc_error(self.h)

    def __getattr__(self, name):
        if name.startswith("_") or name in ("L", "h", "case", "ids"):
            raise AttributeError(name)
        fn = getattr(self.L, "orc_" + name)
        h = self.h
        return lambda *a: fn(h, *a)


class Adjoint:
    """The oracle's discrete adjoint (liboracle_adj.so): the oracle's own arithmetic recorded on a tape and
    swept backwards.  set_initial(f, x0) makes every element of x0 an input (set into M, current and
    forecast, as dyn_adjoint_backward defines an initial field); gradient(seeds) returns dJ/dx0 of
    J = sum(seed * final current field) for u, v, w, theta.  One instance at a time per process (the tape
    is shared)."""

    def __init__(self, case, initial=None):
        self.L = lib("adj")
        self.case = case
        from model_amd import ids
        self.ids = ids
        g = OrcGrid(case.nx, case.ny, case.nz, case.dx, case.dz, case.dt, case.microphysics,
                    case.rayleigh_level, case.sponge, case.sponge)
        self.h = self.L.orc_adj_create(ctypes.byref(g))
        for name, arr in case.columns.items():
            v = np.ascontiguousarray(arr, dtype=np.float64)
            self.L.orc_adj_set_column(self.h, ids.COLUMN_ID[name], _ptr(v))
        for name, arr in case.fields.items():
            f = ids.FIELD_ID[name]
            v = np.ascontiguousarray(arr, dtype=np.float64)
            for role in ((ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P) if f in ids.PROGNOSTIC else (0,)):
                self.L.orc_adj_set_field(self.h, f, role, _ptr(v))
        if initial is not None:
            for f, x0 in zip((ids.U, ids.V, ids.W, ids.TH), initial):
                self.set_initial(f, x0)

    def close(self):
        if getattr(self, "h", None):
            self.L.orc_adj_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def set_initial(self, f, value):
        v = np.ascontiguousarray(value, dtype=np.float64)
        assert self.L.orc_adj_set_initial(self.h, f, _ptr(v)) == 0

    def step(self, n=1):
        self.L.orc_adj_step(self.h, n)

    def get_field(self, f, role=1):
        c = self.case
        out = np.empty((c.nx, c.ny, c.nz))
        self.L.orc_adj_get_field(self.h, f, role, _ptr(out))
        return out

    @property
    def tape_size(self):
        return self.L.orc_adj_tape_size()

    def gradient(self, seeds):
        """seeds: [u, v, w, theta] arrays or None; returns the four gradients dJ/dx0."""
        c = self.case
        sp = [None if sd is None else np.ascontiguousarray(sd, dtype=np.float64) for sd in seeds]
        grads = [np.zeros((c.nx, c.ny, c.nz)) for _ in range(4)]
        rc = self.L.orc_adj_gradient(self.h, *[None if sd is None else _ptr(sd) for sd in sp], *[_ptr(g) for g in grads])
        assert rc == 0
        return grads


class Tangent:
    """The oracle's tangent-linear model (liboracle_tlm.so): every field carries a tangent, and a step
    advances both, differentiating the oracle's own arithmetic with its branches frozen at the base
    state.  Fields are set with an optional tangent; get_field returns (value, tangent)."""

    def __init__(self, case):
        self.L = lib("tlm")
        self.case = case
        g = OrcGrid(case.nx, case.ny, case.nz, case.dx, case.dz, case.dt, case.microphysics,
                    case.rayleigh_level, case.sponge, case.sponge)
        self.h = self.L.orc_tlm_create(ctypes.byref(g))
        from model_amd import ids
        self.ids = ids
        for name, arr in case.columns.items():
            v = np.ascontiguousarray(arr, dtype=np.float64)
            self.L.orc_tlm_set_column(self.h, ids.COLUMN_ID[name], _ptr(v))
        for name, arr in case.fields.items():
            f = ids.FIELD_ID[name]
            for role in ((ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P) if f in ids.PROGNOSTIC else (0,)):
                self.set_field(f, role, arr)

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_tlm_destroy(self.h)
            self.h = None

    def set_field(self, f, role, value, tangent=None):
        v = np.ascontiguousarray(value, dtype=np.float64)
        t = None if tangent is None else np.ascontiguousarray(tangent, dtype=np.float64)
        rc = self.L.orc_tlm_set_field(self.h, f, role, _ptr(v), None if t is None else _ptr(t))
        assert rc == 0

    def get_field(self, f, role=1):
        c = self.case
        shape = (c.nx, c.ny) if f == self.ids.ACCRAIN else (c.nx, c.ny, c.nz)
        v, t = np.empty(shape), np.empty(shape)
        assert self.L.orc_tlm_get_field(self.h, f, role, _ptr(v), _ptr(t)) == 0
        return v, t

    def step(self, n=1):
        self.L.orc_tlm_step(self.h, n)
