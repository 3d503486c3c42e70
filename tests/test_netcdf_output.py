"""The reference's NetCDF output file, written by the library's own classic-format writer (CPU).

The file is read back with scipy.io.netcdf_file, an independent reader of the netCDF classic and 64-bit
offset formats: dimensions, global attributes and variables must be those of create_outfile
(files.cpp:51-282), records must hold what was written (as NC_FLOAT), the record count must grow with
the records, and the restart attribute last_file_successfully_written must round-trip
(files.cpp:1205-1254).  The GPU staging path that feeds the records is in tests/test_gpu_output.py.
"""
import ctypes

import numpy as np
import pytest

scipy_netcdf = pytest.importorskip("scipy.io")

from model_amd.core import NcMeta, load  # noqa: E402

NX, NY, NZ = 14, 10, 7


def _lib():
    return load()


def _meta():
    return NcMeta(dx=1000.0, dz=250.0, dt=2.0, latoffset=10.0, lonoffset=20.0, water_temperature=28.0,
                  height_lowest_level=250.0, use_turbulent_stress=0, basic_state_option=1, perturbation_option=1,
                  reanalysis_input_data=b"none", perturbation_input_data=b"bubble", basic_state_input_data=b"")


def _put(L, path, var, rec, a):
    a = np.ascontiguousarray(a, dtype=np.float64)
    rc = L.dyn_netcdf_put(path, var.encode(), rec, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), a.size)
    assert rc == 0, L.dyn_last_error()


@pytest.mark.parametrize("moist", [0, 1])
def test_layout_matches_create_outfile(tmp_path, moist):
    L = _lib()
    path = str(tmp_path / "out.nc").encode()
    m = _meta()
    assert L.dyn_netcdf_create(path, NX, NY, NZ, moist, ctypes.byref(m)) == 0, L.dyn_last_error()
    with scipy_netcdf.netcdf_file(path.decode(), "r", mmap=False) as f:
        assert f.version_byte == 2  # 64-bit offset (NC64BIT, file.h:13)
        assert f.dimensions == {"x": NX, "y": NY, "z": NZ, "t": None}
        names = ["u-wind", "v-wind", "w-wind", "theta", "pi"] + (["rainfall", "qv", "qc", "qr"] if moist else []) + \
            ["ubar", "vbar", "wbar", "thbar", "qbar", "pbar", "tb", "pib", "qb", "zu", "lat", "lon", "time", "topo"]
        assert list(f.variables) == names
        for n in names:
            assert f.variables[n].typecode() == "f", n
        assert f.variables["u-wind"].dimensions == ("t", "x", "y", "z")
        assert f.variables["ubar"].dimensions == ("x", "y", "z")
        assert f.variables["tb"].dimensions == ("z",)
        assert f.variables["lat"].dimensions == ("y",) and f.variables["lon"].dimensions == ("x",)
        assert f.variables["topo"].dimensions == ("x", "y")
        if moist:
            assert f.variables["rainfall"].dimensions == ("t", "x", "y")
        a = f._attributes
        assert float(a["dx"]) == 1000.0 and float(a["dy"]) == 1000.0 and float(a["dz"]) == 250.0
        assert int(a["microphysics_option"]) == moist and int(a["periodic_boundaries"]) == 1
        assert int(a["stretched_grid"]) == 1 and int(a["hydrostatic"]) == 0
        assert int(a["last_file_successfully_written"]) == 2
        assert a["perturbation_input_data"] == b"bubble"
        assert f.variables["u-wind"].shape[0] == 0  # no records yet


def test_records_and_time_round_trip(tmp_path):
    L = _lib()
    path = str(tmp_path / "out.nc").encode()
    m = _meta()
    assert L.dyn_netcdf_create(path, NX, NY, NZ, 1, ctypes.byref(m)) == 0
    rng = np.random.default_rng(3)
    data = {}
    for rec in range(3):
        for v in ("u-wind", "v-wind", "w-wind", "theta", "pi", "qv", "qc", "qr"):
            a = rng.standard_normal((NX, NY, NZ)) * 10.0 ** rng.integers(-5, 3)
            data[(v, rec)] = a
            _put(L, path, v, rec, a)
        r2 = rng.random((NX, NY))
        data[("rainfall", rec)] = r2
        _put(L, path, "rainfall", rec, r2)
        _put(L, path, "time", rec, np.array([0.5 * rec]))
        assert L.dyn_netcdf_records(path) == rec + 1
    zu = np.linspace(-125, 1500, NZ)
    _put(L, path, "zu", 0, zu)
    ub = rng.standard_normal((NX, NY, NZ))
    _put(L, path, "ubar", 0, ub)
    with scipy_netcdf.netcdf_file(path.decode(), "r", mmap=False) as f:
        assert f.variables["u-wind"].shape == (3, NX, NY, NZ)
        for (v, rec), a in data.items():
            np.testing.assert_array_equal(f.variables[v][rec], a.astype(np.float32))
        np.testing.assert_array_equal(f.variables["time"][:], np.array([0, 0.5, 1.0], dtype=np.float32))
        np.testing.assert_array_equal(f.variables["zu"][:], zu.astype(np.float32))
        np.testing.assert_array_equal(f.variables["ubar"][:], ub.astype(np.float32))
        assert not f.variables["topo"][:].any()
    # the library's own reader agrees
    out = np.empty((NX, NY, NZ))
    assert L.dyn_netcdf_read(path, b"theta", 1, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size) == 0
    np.testing.assert_array_equal(out, data[("theta", 1)].astype(np.float32).astype(np.float64))
    assert L.dyn_netcdf_read(path, b"theta", 3, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), out.size) == -1


def test_restart_status_attribute(tmp_path):
    L = _lib()
    path = str(tmp_path / "out.nc").encode()
    m = _meta()
    assert L.dyn_netcdf_create(path, NX, NY, NZ, 0, ctypes.byref(m)) == 0
    st = ctypes.c_int()
    assert L.dyn_netcdf_get_status(path, ctypes.byref(st)) == 0 and st.value == 2
    for value in (0, 1, 2):
        assert L.dyn_netcdf_set_status(path, value) == 0
        assert L.dyn_netcdf_get_status(path, ctypes.byref(st)) == 0 and st.value == value
        with scipy_netcdf.netcdf_file(path.decode(), "r", mmap=False) as f:
            assert int(f._attributes["last_file_successfully_written"]) == value


def test_status_of_a_file_without_the_attribute(tmp_path):
    """get_file_output_status returns -1 when the attribute is absent (files.cpp:1232-1254); a file
    written by another classic-format writer is read as well."""
    L = _lib()
    p = tmp_path / "other.nc"
    with scipy_netcdf.netcdf_file(str(p), "w", version=2) as f:
        f.createDimension("z", 4)
        v = f.createVariable("zu", "f", ("z",))
        v[:] = np.arange(4, dtype=np.float32)
    st = ctypes.c_int(7)
    assert L.dyn_netcdf_get_status(str(p).encode(), ctypes.byref(st)) == 0 and st.value == -1
    assert L.dyn_netcdf_set_status(str(p).encode(), 1) == -2
    out = np.empty(4)
    assert L.dyn_netcdf_read(str(p).encode(), b"zu", 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 4) == 0
    np.testing.assert_array_equal(out, np.arange(4.0))
