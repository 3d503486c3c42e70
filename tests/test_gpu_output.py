"""Output staging, the NetCDF record path and failure detection on the device (needs an MI355X).

- dyn_output_stage snapshots the output fields (u, v, w, pi, theta; qv, qc, qr, rainfall) and copies them
  to the host on a stream of its own while later steps run; the staged values must be the fields at the
  time of the call, narrowed to float, whatever the steps queued after it do.
- dyn_netcdf_init / dyn_netcdf_write_staged write the reference's file (outfile_init,
  output_meteorological_fields_to_file, write_time_to_file); read back with scipy's independent reader.
- dyn_check_finite reports a NaN uploaded into a field, and the Kessler guard E > 100 (kessler.cpp:83-88)
  is raised through dyn_synchronize by an absurd rain mixing ratio: both forced from the host with data.
"""
import ctypes

import numpy as np
import pytest

from helpers import zero_case
from model_amd import configs, ids
from model_amd.core import DYN_FP32, DynCore, DynError, NcMeta

pytestmark = pytest.mark.gpu

OUT = [(ids.U, "u-wind"), (ids.V, "v-wind"), (ids.W, "w-wind"), (ids.PI, "pi"), (ids.TH, "theta"),
       (ids.QV, "qv"), (ids.QC, "qc"), (ids.QR, "qr"), (ids.ACCRAIN, "rainfall")]


def _staged(g, f):
    c = g.case
    shape = (c.nx, c.ny) if f == ids.ACCRAIN else (c.nx, c.ny, c.nz)
    out = np.empty(shape, dtype=np.float32)
    rc = g.L.dyn_output_get(g.h, f, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)), out.size)
    assert rc == 0, g.L.dyn_last_error()
    return out


@pytest.mark.parametrize("precision", [0, DYN_FP32], ids=["fp64", "fp32"])
def test_staged_fields_are_the_fields_at_the_call(precision):
    c = configs.supercell(nx=48, ny=40, nz=30)
    g = DynCore(c, precision=precision)
    g.step(5)
    g.sync()
    ref = {f: g.get_field(f, 0 if f in (ids.PI, ids.ACCRAIN) else ids.ROLE_CUR) for f, _ in OUT}
    assert g.L.dyn_output_stage(g.h) == 0, g.L.dyn_last_error()
    g.step(7)  # queued behind the snapshot; they must not change what was staged
    for f, _ in OUT:
        np.testing.assert_array_equal(_staged(g, f), ref[f].astype(np.float32))
    assert g.L.dyn_output_ready(g.h) == 1
    # a second stage after more steps sees the new state
    g.sync()
    later = g.get_field(ids.TH, ids.ROLE_CUR)
    assert g.L.dyn_output_stage(g.h) == 0
    np.testing.assert_array_equal(_staged(g, ids.TH), later.astype(np.float32))
    g.close()


def test_netcdf_records_from_staged_output(tmp_path):
    scipy_io = pytest.importorskip("scipy.io")
    c = configs.supercell(nx=48, ny=40, nz=30)
    g = DynCore(c, precision=DYN_FP32)
    path = str(tmp_path / "run.nc").encode()
    meta = NcMeta(dx=c.dx, dz=c.dz, dt=c.dt, height_lowest_level=c.dz, basic_state_option=1, perturbation_option=1)
    lat = np.linspace(10, 12, c.ny)
    lon = np.linspace(80, 83, c.nx)
    dp = ctypes.POINTER(ctypes.c_double)
    assert g.L.dyn_netcdf_init(g.h, path, ctypes.byref(meta), lat.ctypes.data_as(dp), lon.ctypes.data_as(dp)) == 0, \
        g.L.dyn_last_error()
    expect = []
    for rec in range(3):  # s_run_model: output, then the steps up to the next output
        g.sync()
        expect.append({f: g.get_field(f, 0 if f in (ids.PI, ids.ACCRAIN) else ids.ROLE_CUR) for f, _ in OUT})
        assert g.L.dyn_output_stage(g.h) == 0
        g.step(10)
        assert g.L.dyn_netcdf_write_staged(g.h, path, rec, rec * 10 * c.dt / 3600.0) == 0, g.L.dyn_last_error()
    assert g.L.dyn_netcdf_set_status(path, 2) == 0
    with scipy_io.netcdf_file(path.decode(), "r", mmap=False) as fh:
        v = fh.variables
        assert v["theta"].shape == (3, c.nx, c.ny, c.nz)
        for rec in range(3):
            for f, name in OUT:
                np.testing.assert_array_equal(v[name][rec], expect[rec][f].astype(np.float32), err_msg=f"{name} {rec}")
        np.testing.assert_allclose(v["time"][:], [0.0, 20 / 3600.0, 40 / 3600.0], rtol=1e-6)
        np.testing.assert_array_equal(v["ubar"][:], c.fields["ubar"].astype(np.float32))
        np.testing.assert_array_equal(v["tb"][:], c.columns["tb"].astype(np.float32))
        np.testing.assert_array_equal(v["zu"][:], c.columns["zsu"].astype(np.float32))
        np.testing.assert_array_equal(v["lat"][:], lat.astype(np.float32))
        assert int(fh._attributes["last_file_successfully_written"]) == 2
    g.close()


@pytest.mark.parametrize("precision", [0, DYN_FP32], ids=["fp64", "fp32"])
def test_check_finite_reports_a_nan(precision):
    c = zero_case(nx=16, ny=16, nz=12, microphysics=1)
    g = DynCore(c, precision=precision)
    mask = ctypes.c_int(-1)
    assert g.L.dyn_check_finite(g.h, ctypes.byref(mask)) == 0 and mask.value == 0
    th = np.zeros((c.nx, c.ny, c.nz))
    th[5, 6, 7] = np.nan
    g.set_field(ids.TH, ids.ROLE_CUR, th)
    qr = np.zeros((c.nx, c.ny, c.nz))
    qr[1, 2, 3] = np.inf
    g.set_field(ids.QR, ids.ROLE_CUR, qr)
    assert g.L.dyn_check_finite(g.h, ctypes.byref(mask)) == -4
    assert mask.value == (1 << ids.TH) | (1 << ids.QR)
    assert b"theta" in g.L.dyn_last_error() and b"qr" in g.L.dyn_last_error()
    g.close()


def test_kessler_guard_raises_numeric_error():
    """kessler.cpp:83-88 aborts when the evaporation E exceeds 100; here an absurd rain mixing ratio in
    a subsaturated column makes E > 100, and dyn_synchronize returns DYN_E_NUMERIC."""
    c = zero_case(nx=16, ny=16, nz=12, microphysics=1)
    g = DynCore(c)
    qr = np.zeros((c.nx, c.ny, c.nz))
    qr[8, 8, 3] = 1.0e8
    g.set_field(ids.QR, ids.ROLE_P, qr)
    qv = np.full((c.nx, c.ny, c.nz), -0.5 * c.columns["qb"][3])  # well below saturation
    g.set_field(ids.QV, ids.ROLE_P, qv)
    g.kessler(1, c.nx - 1, 1, c.ny - 1)
    with pytest.raises(DynError) as e:
        g.sync()
    assert e.value.code == -4 and "E > 100" in str(e.value)
    g.close()
