"""The KPPHvalue restart variable (gckpp_kpph_restart_write / _read), on the
CPU, against an independent netCDF classic implementation (scipy.io's
netcdf_file): what the library writes, scipy reads as Chem_KPPHvalue(time,
lev, lat, lon) with the same values, and what scipy writes (CDF-1 and
CDF-2, double and float) the library reads back exactly; a file without
the variable gives zeros and found = false, the reference's cold start
(GeosCore/hco_utilities_gc_mod.F90:2144-2177)."""
import numpy as np
import pytest


@pytest.fixture(scope="module")
def hip_lib():
    from gckpp_amd import hip
    if not hip.lib_path().exists():
        pytest.skip("HIP library not built")
    hip.load()
    return hip


def _field(shape, seed):
    rng = np.random.default_rng(seed)
    k = 10.0 ** rng.uniform(-5, np.log10(1200.0), shape)
    k[0, 0, 0] = 0.0
    return k


def test_written_file_reads_in_scipy(hip_lib, tmp_path):
    from scipy.io import netcdf_file
    hip = hip_lib
    k = _field((72, 46, 72), 1)
    p = tmp_path / "restart.nc"
    hip.kpph_restart_write(p, k, model_tag="h0_mlp_fullchem v1")
    with netcdf_file(str(p), "r", mmap=False) as f:
        v = f.variables["Chem_KPPHvalue"]
        assert v.dimensions == ("time", "lev", "lat", "lon")
        assert v.data.shape == (1, 72, 46, 72)
        assert v.units == b"s"
        assert f.gckpp_step_model == b"h0_mlp_fullchem v1"
        # Fortran KPPHvalue(lon, lat, lev) is the C array [time][lev][lat][lon]
        assert np.array_equal(np.asarray(v.data[0]).transpose(2, 1, 0), k)
    k2, found, tag = hip.kpph_restart_read(p, k.shape)
    assert found and tag == "h0_mlp_fullchem v1" and np.array_equal(k2, k)


@pytest.mark.parametrize("version,dtype", [(1, "f8"), (2, "f8"), (1, "f4")])
def test_reads_scipy_written_files(version, dtype, hip_lib, tmp_path):
    from scipy.io import netcdf_file
    hip = hip_lib
    k = _field((12, 7, 5), 2)
    p = tmp_path / "scipy.nc"
    with netcdf_file(str(p), "w", version=version) as f:
        f.createDimension("time", None)             # a record dimension, as restart files have
        f.createDimension("lev", 5)
        f.createDimension("lat", 7)
        f.createDimension("lon", 12)
        o = f.createVariable("SpeciesRst_O3", "f8", ("time", "lev", "lat", "lon"))
        o[0] = np.ones((5, 7, 12))
        v = f.createVariable("Chem_KPPHvalue", dtype, ("time", "lev", "lat", "lon"))
        v.units = "s"
        v[0] = k.transpose(2, 1, 0).astype(dtype)
    k2, found, tag = hip.kpph_restart_read(p, k.shape)
    assert found and tag == ""
    assert np.array_equal(k2, k.astype(dtype).astype(np.float64))


def test_missing_variable_is_a_cold_start(hip_lib, tmp_path):
    from scipy.io import netcdf_file
    hip = hip_lib
    p = tmp_path / "novar.nc"
    with netcdf_file(str(p), "w") as f:
        f.createDimension("lon", 4)
        f.createVariable("other", "f8", ("lon",))[:] = 1.0
    k2, found, _ = hip.kpph_restart_read(p, (4, 1, 1))
    assert not found and np.array_equal(k2, np.zeros((4, 1, 1)))
    with pytest.raises(hip.GckppError, match="grid's size"):
        q = tmp_path / "small.nc"
        hip.kpph_restart_write(q, _field((3, 2, 2), 3))
        hip.kpph_restart_read(q, (4, 2, 2))
    bad = tmp_path / "bad.nc"
    bad.write_bytes(b"HDF\x89 not netcdf")
    with pytest.raises(hip.GckppError, match="not a netCDF classic"):
        hip.kpph_restart_read(bad, (4, 1, 1))
