"""Python host binding of libdyncore's C ABI (include/dyncore.h) through ctypes.

The library is loaded from model_amd/_build/libdyncore.so; if it is missing the import fails loudly --
there is no CPU fallback.  Host arrays are float64 numpy arrays in the reference layout [i][j][k].
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import ids

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libdyncore.so")
REFAPI_PATH = os.path.join(_HERE, "_build", "libdyncore_refapi.so")

DYN_FP64, DYN_FP32 = 0, 1
DYN_TUNE_MOMENTUM_TJ, DYN_TUNE_MOMENTUM_TI, DYN_TUNE_SCALAR_TJ, DYN_TUNE_SCALAR_TI, DYN_TUNE_PRESSURE_JPASS = 0, 1, 2, 3, 4
DYN_TUNE_STAGE_STREAMS = 5
ERRORS = {-1: "DYN_E_ARG", -2: "DYN_E_UNSUPPORTED", -3: "DYN_E_HIP", -4: "DYN_E_NUMERIC", -5: "DYN_E_STATE"}

# the exported C symbols, as declared in include/dyncore.h
EXPORTS = [
    "dyn_last_error", "dyn_version", "dyn_create", "dyn_destroy", "dyn_set_column", "dyn_set_field", "dyn_get_field",
    "dyn_synchronize", "dyn_advect_uvw", "dyn_advect_theta", "dyn_advect_moisture", "dyn_zero_moisture",
    "dyn_solve_pressure", "dyn_pgf", "dyn_apply_bc", "dyn_apply_bc_micro", "dyn_kessler", "dyn_rayleigh",
    "dyn_advance_inner", "dyn_advance_outer", "dyn_micro_advance_inner", "dyn_micro_advance_outer",
    "dyn_step_reference", "dyn_step", "dyn_step_fused", "dyn_algorithmic_bytes_per_point", "dyn_timing", "dyn_timing_report",
    "dyn_stream", "dyn_tune_set", "dyn_tune_get", "dyn_adjoint_record", "dyn_adjoint_backward", "dyn_adjoint_release",
    "dyn_adjoint_interval",
    "dyn_ensemble_create", "dyn_ensemble_destroy", "dyn_ensemble_synchronize", "dyn_ensemble_step", "dyn_ensemble_step_eager",
    "dyn_ensemble_set_shape", "dyn_ensemble_autotune", "dyn_ensemble_set_groups", "dyn_ensemble_groups",
    "dyn_get_field_noalloc", "dyn_check_finite", "dyn_output_stage", "dyn_output_ready", "dyn_output_get", "dyn_netcdf_create", "dyn_netcdf_init",
    "dyn_netcdf_write_staged", "dyn_netcdf_put", "dyn_netcdf_read", "dyn_netcdf_records", "dyn_netcdf_set_status",
    "dyn_netcdf_get_status", "dyn_set_turbulence", "dyn_calculate_diff_tend", "dyn_apply_velocity_diffusion",
    "dyn_apply_scalar_diffusion", "dyn_apply_moisture_diffusion", "dyn_get_turbulence",
]


class DynError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"{ERRORS.get(code, code)}: {message}")
        self.code = code


class DynGrid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
                ("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
                ("microphysics", ctypes.c_int), ("periodic_ew", ctypes.c_int), ("rayleigh_level", ctypes.c_int),
                ("sponge_north", ctypes.c_int), ("sponge_south", ctypes.c_int), ("precision", ctypes.c_int)]


class NcMeta(ctypes.Structure):
    """dyn_nc_meta: the global attributes of the output file that the configuration sets."""
    _fields_ = [("dx", ctypes.c_double), ("dz", ctypes.c_double), ("dt", ctypes.c_double),
                ("latoffset", ctypes.c_double), ("lonoffset", ctypes.c_double), ("water_temperature", ctypes.c_double),
                ("height_lowest_level", ctypes.c_double), ("use_turbulent_stress", ctypes.c_int),
                ("basic_state_option", ctypes.c_int), ("perturbation_option", ctypes.c_int),
                ("reanalysis_input_data", ctypes.c_char_p), ("perturbation_input_data", ctypes.c_char_p),
                ("basic_state_input_data", ctypes.c_char_p)]


_lib = None
_dp = ctypes.POINTER(ctypes.c_double)


def load():
    """Load libdyncore.so (built in-tree by model_amd.build)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    L = ctypes.CDLL(LIB_PATH)
    L.dyn_last_error.restype = ctypes.c_char_p
    L.dyn_create.argtypes = [ctypes.POINTER(DynGrid), ctypes.POINTER(ctypes.c_void_p)]
    L.dyn_destroy.argtypes = [ctypes.c_void_p]
    L.dyn_set_column.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_set_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_get_field.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_synchronize.argtypes = [ctypes.c_void_p]
    for fn in ("dyn_advect_uvw", "dyn_advect_theta", "dyn_advect_moisture", "dyn_pgf"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4
    for fn in ("dyn_zero_moisture", "dyn_kessler"):
        getattr(L, fn).argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
    L.dyn_rayleigh.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 5
    L.dyn_solve_pressure.argtypes = [ctypes.c_void_p, ctypes.c_double]
    for fn in ("dyn_apply_bc", "dyn_apply_bc_micro", "dyn_advance_inner", "dyn_advance_outer", "dyn_micro_advance_inner",
               "dyn_micro_advance_outer"):
        getattr(L, fn).argtypes = [ctypes.c_void_p]
    L.dyn_step_reference.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_step_fused.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_algorithmic_bytes_per_point.argtypes = [ctypes.c_void_p]
    L.dyn_algorithmic_bytes_per_point.restype = ctypes.c_double
    L.dyn_timing.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_timing_report.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t]
    L.dyn_stream.argtypes = [ctypes.c_void_p]
    L.dyn_tune_set.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.dyn_tune_get.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_adjoint_record.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    L.dyn_adjoint_backward.argtypes = [ctypes.c_void_p] + [_dp] * 8
    L.dyn_adjoint_release.argtypes = [ctypes.c_void_p]
    L.dyn_adjoint_interval.argtypes = [ctypes.c_void_p]
    L.dyn_ensemble_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]
    for fn in ("dyn_ensemble_destroy", "dyn_ensemble_synchronize"):
        getattr(L, fn).argtypes = [ctypes.c_void_p]
    for fn in ("dyn_ensemble_step", "dyn_ensemble_step_eager"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_ensemble_set_shape.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.dyn_ensemble_autotune.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    L.dyn_ensemble_set_groups.argtypes = [ctypes.c_void_p, ctypes.c_int]
    L.dyn_ensemble_groups.argtypes = [ctypes.c_void_p]
    L.dyn_stream.restype = ctypes.c_void_p
    L.dyn_get_field_noalloc.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_check_finite.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
    for fn in ("dyn_output_stage", "dyn_output_ready"):
        getattr(L, fn).argtypes = [ctypes.c_void_p]
    L.dyn_output_get.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_float), ctypes.c_size_t]
    L.dyn_netcdf_create.argtypes = [ctypes.c_char_p] + [ctypes.c_int] * 4 + [ctypes.POINTER(NcMeta)]
    L.dyn_netcdf_init.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(NcMeta), _dp, _dp]
    L.dyn_netcdf_write_staged.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_double]
    for fn in ("dyn_netcdf_put", "dyn_netcdf_read"):
        getattr(L, fn).argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_netcdf_records.argtypes = [ctypes.c_char_p]
    L.dyn_netcdf_set_status.argtypes = [ctypes.c_char_p, ctypes.c_int]
    L.dyn_netcdf_get_status.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_int)]
    L.dyn_set_turbulence.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t]
    L.dyn_calculate_diff_tend.argtypes = [ctypes.c_void_p] + [ctypes.c_int] * 4
    for fn in ("dyn_apply_velocity_diffusion", "dyn_apply_scalar_diffusion", "dyn_apply_moisture_diffusion"):
        getattr(L, fn).argtypes = [ctypes.c_void_p, ctypes.c_double] + [ctypes.c_int] * 4
    L.dyn_get_turbulence.argtypes = [ctypes.c_void_p, ctypes.c_int, _dp, ctypes.c_size_t]
    _lib = L
    return L


def _check(rc: int):
    if rc != 0:
        raise DynError(rc, load().dyn_last_error().decode())


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_dp)


class DynCore:
    """One HBM-resident model state (a dyn_ctx)."""

    def __init__(self, case, precision: int = DYN_FP64):
        self.L = load()
        self.case = case
        g = DynGrid(case.nx, case.ny, case.nz, case.dx, case.dz, case.dt, case.microphysics, 1, case.rayleigh_level,
                    case.sponge, case.sponge, precision)
        h = ctypes.c_void_p()
        _check(self.L.dyn_create(ctypes.byref(g), ctypes.byref(h)))
        self.h = h
        self.precision = precision
        for name, arr in case.columns.items():
            self.set_column(ids.COLUMN_ID[name], arr)
        for name, arr in case.fields.items():
            f = ids.FIELD_ID[name]
            if f in (ids.QV, ids.QC, ids.QR) and not case.microphysics:
                continue
            if f in ids.PROGNOSTIC:
                for role in (ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P):
                    self.set_field(f, role, arr)
            else:
                self.set_field(f, 0, arr)

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.dyn_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_column(self, c, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.L.dyn_set_column(self.h, c, _ptr(v), v.size))

    def set_field(self, f, role, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        _check(self.L.dyn_set_field(self.h, f, role, _ptr(v), v.size))

    def get_field(self, f, role=ids.ROLE_CUR):
        c = self.case
        shape = (c.nx, c.ny) if f == ids.ACCRAIN else (c.nx, c.ny, c.nz)
        out = np.empty(shape)
        _check(self.L.dyn_get_field(self.h, f, role, _ptr(out), out.size))
        return out

    def sync(self):
        _check(self.L.dyn_synchronize(self.h))

    def set_turbulence(self, on=True, landsea=None):
        """USE_TURBULENT_STRESS on or off; landsea: NX x NY mask (< 0.5 land) or None for all ocean."""
        if landsea is None:
            _check(self.L.dyn_set_turbulence(self.h, int(on), None, 0))
        else:
            ls = np.ascontiguousarray(landsea, dtype=np.float64)
            _check(self.L.dyn_set_turbulence(self.h, int(on), _ptr(ls), ls.size))

    def get_turbulence(self, which):
        """0 Kmix, 1 KHmix, 2-4 u/v/w friction, 5-8 theta/qv/qc/qr diffusion tendencies, 9 saturation flag."""
        c = self.case
        out = np.empty((c.nx, c.ny, c.nz))
        _check(self.L.dyn_get_turbulence(self.h, which, _ptr(out), out.size))
        return out

    def step(self, n=1):
        _check(self.L.dyn_step(self.h, n))

    def step_reference(self, n=1):
        _check(self.L.dyn_step_reference(self.h, n))

    def step_fused(self, n=1):
        _check(self.L.dyn_step_fused(self.h, n))

    @property
    def stream(self):
        return self.L.dyn_stream(self.h)

    def bytes_per_point(self):
        return self.L.dyn_algorithmic_bytes_per_point(self.h)

    def tune_get(self, param: int) -> int:
        v = self.L.dyn_tune_get(self.h, param)
        _check(min(v, 0))
        return v

    def __getattr__(self, name):
        if name.startswith("_") or name in ("L", "h", "case", "ids"):
            raise AttributeError(name)
        fn = getattr(self.L, "dyn_" + name)
        h = self.h

        def call(*a):
            _check(fn(h, *a))

        return call


class Ensemble:
    """Members of possibly different horizontal sizes advanced together (dyn_ensemble_*)."""

    def __init__(self, members):
        self.L = load()
        self.members = list(members)
        arr = (ctypes.c_void_p * len(self.members))(*[m.h.value for m in self.members])
        h = ctypes.c_void_p()
        _check(self.L.dyn_ensemble_create(arr, len(self.members), ctypes.byref(h)))
        self.h = h

    def step(self, n=1, graph=True):
        _check((self.L.dyn_ensemble_step if graph else self.L.dyn_ensemble_step_eager)(self.h, n))

    def sync(self):
        _check(self.L.dyn_ensemble_synchronize(self.h))

    def set_shape(self, shape):
        arr = (ctypes.c_int * 4)(*shape)
        _check(self.L.dyn_ensemble_set_shape(self.h, arr))

    def autotune(self):
        arr = (ctypes.c_int * 4)()
        _check(self.L.dyn_ensemble_autotune(self.h, arr))
        return tuple(arr)

    def set_groups(self, n):
        _check(self.L.dyn_ensemble_set_groups(self.h, n))

    @property
    def groups(self):
        n = self.L.dyn_ensemble_groups(self.h)
        _check(min(n, 0))
        return n

    def close(self):
        if getattr(self, "h", None) and self.h.value:
            self.L.dyn_ensemble_destroy(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
