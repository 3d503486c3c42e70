"""dyn_get_field_noalloc: the read-back entry point that allocates nothing (needs an MI355X).

It must return what dyn_get_field returns, bit for bit, on two calls in a row with the second on a
smaller field (the shared device staging array keeps the first call's contents), in FP32 and FP64, and
reject a field larger than its staging array.
"""
import ctypes

import numpy as np
import pytest

from model_amd import configs, ids
from model_amd.core import DYN_FP32, DynCore, load

pytestmark = pytest.mark.gpu

_dp = ctypes.POINTER(ctypes.c_double)


def _noalloc():
    L = load()
    fn = L.dyn_get_field_noalloc
    fn.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_size_t]
    return L, fn


def _read(fn, core, f, role):
    c = core.case
    shape = (c.nx, c.ny) if f == ids.ACCRAIN else (c.nx, c.ny, c.nz)
    out = np.full(shape, np.nan)
    rc = fn(core.h, f, role, out.ctypes.data_as(_dp), out.size)
    assert rc == 0, load().dyn_last_error()
    return out


@pytest.mark.parametrize("precision", [DYN_FP32, 0], ids=["fp32", "fp64"])
def test_noalloc_read_back_matches_get_field(precision):
    L, fn = _noalloc()
    big = DynCore(configs.supercell(), precision=precision)   # 262 x 256 x 64 = 4,292,608 values per field
    big.step(3)
    big.sync()
    small = DynCore(configs.warm_bubble(nx=32, ny=32, nz=20), precision=precision)
    small.step(2)
    small.sync()
    for core, fields in ((big, (ids.U, ids.W, ids.TH, ids.QR, ids.PI, ids.ACCRAIN)), (small, (ids.V, ids.TH))):
        for f in fields:
            for role in ((ids.ROLE_M, ids.ROLE_CUR, ids.ROLE_P) if f in ids.PROGNOSTIC else (0,)):
                ref = core.get_field(f, role)
                got = _read(fn, core, f, role)
                assert np.array_equal(ref, got), (core.case.name, ids.FIELD_NAMES[f], role)
    big.close()
    small.close()


def test_noalloc_rejects_a_field_beyond_its_staging_array():
    L, fn = _noalloc()
    c = configs.supercell(nx=256, ny=256, nz=130, bubble=False)   # 8,719,360 values > 8,388,608
    core = DynCore(c, precision=DYN_FP32)
    out = np.empty((c.nx, c.ny, c.nz))
    rc = fn(core.h, ids.TH, ids.ROLE_CUR, out.ctypes.data_as(_dp), out.size)
    assert rc == -1
    assert b"staging array" in L.dyn_last_error()
    # the allocating read-back still works, and a later small call through the staging array is exact
    ref = core.get_field(ids.TH, ids.ROLE_CUR)
    assert np.isfinite(ref).all()
    core.close()
    small = DynCore(configs.warm_bubble(nx=32, ny=32, nz=20), precision=DYN_FP32)
    assert np.array_equal(small.get_field(ids.TH, ids.ROLE_CUR), _read(fn, small, ids.TH, ids.ROLE_CUR))
    small.close()
