"""KPP-Standalone chemical-state files (gckpp_kppsa_write / _read), on the
CPU: the layout of GeosCore/kppsa_interface_mod.F90:569-843 (a 60-line
header whose first line is that count, the metadata lines at the positions
the reference's WRITE sequence puts them, the column line, then NSPEC
species, NREACT R and NREACT A lines), every number exactly as the Fortran
runtime renders the reference's edit descriptors (oracle/kppsa_fmt, built
by flang, formats the same values), and an exact round trip of the values
(es25.16e3 carries 17 significant digits)."""
import struct
import subprocess
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
FMT = ROOT / "oracle" / "_build" / "kppsa_fmt"


@pytest.fixture(scope="module")
def hip_lib():
    from gckpp_amd import hip
    if not hip.lib_path().exists():
        pytest.skip("HIP library not built")
    hip.load()
    return hip


def _meta(hip, **kw):
    m = hip.KppsaMeta()
    m.location = b"Test_Site"
    m.timestamp = b"2019/07/01 12:20"
    m.lon, m.lat, m.level = -117.5, 34.0, 3
    m.pressure_hpa, m.temperature, m.airnumden = 950.25, 288.15, 2.4e19
    m.h2o_vmr, m.cloud_frac, m.suncos = 1.2e-2, 0.0, 0.75
    m.h_init, m.h_exit, m.dt, m.nsteps = 120.5, 300.0, 1200.0, 17
    for k in range(20):
        m.icntrl[k] = 0
        m.rcntrl[k] = 0.0
    m.icntrl[0], m.icntrl[2], m.icntrl[6], m.icntrl[14] = 1, 4, 1, -1
    m.rcntrl[2] = 120.5
    for k, v in kw.items():
        setattr(m, k, v)
    return m


def _values(mech, seed):
    rng = np.random.default_rng(seed)
    conc = 10.0 ** rng.uniform(-20, 19, mech.nspec)
    conc[::7] *= -1.0
    conc[3] = 0.0
    conc[5] = 5e-324                 # subnormal
    conc[6] = 1.7976931348623157e308
    rct = 10.0 ** rng.uniform(-300, 300, mech.nreact)
    rct[::5] = 0.0
    aout = rng.standard_normal(mech.nreact) * 10.0 ** rng.uniform(-30, 30, mech.nreact)
    atol = np.where(np.arange(mech.nspec) % 11 == 0, 1e25, 1e-2)
    return conc, atol, rct, aout


@pytest.mark.parametrize("name", ["fullchem", "Hg", "carbon"])
def test_round_trip_is_exact(name, hip_lib, mech_of, tmp_path):
    hip = hip_lib
    m = mech_of(name)
    conc, atol, rct, aout = _values(m, 1)
    names = list(m.species)
    names[2] = None                       # written as C3
    meta = _meta(hip)
    p = tmp_path / "cell.txt"
    hip.kppsa_write(m, p, meta, conc, atol, rct, aout, names=names)
    meta2, names2, c2, a2, r2, o2 = hip.kppsa_read(m, p)
    assert np.array_equal(c2, conc) and np.array_equal(r2, rct) and np.array_equal(o2, aout)
    assert np.array_equal(a2, atol)
    assert names2[2] == "C3" and names2[:2] == m.species[:2] and names2[3:] == m.species[3:]
    assert meta2.location == b"Test_Site" and meta2.timestamp == b"2019/07/01 12:20"
    assert (meta2.lon, meta2.lat, meta2.level, meta2.nsteps) == (-117.5, 34.0, 3, 17)
    assert meta2.temperature == 288.15 and meta2.pressure_hpa == 950.25
    assert meta2.airnumden == 2.4e19 and meta2.suncos == 0.75
    assert (meta2.h_init, meta2.h_exit, meta2.dt) == (120.5, 300.0, 1200.0)
    assert list(meta2.icntrl) == list(meta.icntrl) and list(meta2.rcntrl) == list(meta.rcntrl)


def test_layout_follows_the_reference_write_sequence(hip_lib, fullchem, tmp_path):
    hip = hip_lib
    conc, atol, rct, aout = _values(fullchem, 2)
    p = tmp_path / "cell.txt"
    hip.kppsa_write(fullchem, p, _meta(hip), conc, atol, rct, aout, names=fullchem.species)
    L = p.read_text().split("\n")
    assert L[-1] == ""
    L = L[:-1]
    assert L[0] == "60"
    assert len(L) == 61 + fullchem.nspec + 2 * fullchem.nreact
    # 1-based line numbers of the reference's WRITE statements
    at = {27: "Meteorological and general grid cell metadata", 29: "Location:", 30: "Timestamp:",
          31: "Longitude (degrees):", 32: "Latitude (degrees):", 33: "GEOS-Chem Vertical Level:",
          34: "Pressure (hPa):", 35: "Temperature (K):", 36: "Dry air density (molec/cm3):",
          37: "Water vapor mixing ratio (vol H2O/vol dry air):", 38: "Cloud fraction:",
          39: "Cosine of solar zenith angle:", 41: "KPP Integrator-specific parameters",
          43: "Init KPP Timestep (seconds):", 44: "Exit KPP Timestep (seconds):",
          45: "Chemistry operator timestep (seconds):", 46: "Number of internal timesteps:",
          47: "ICNTRL integrator options used:", 50: "RCNTRL integrator options used:",
          56: "CSV data of full chemical state", 61: "Name,   Value,   Absolute Tolerance"}
    for n, key in at.items():
        assert L[n - 1].startswith(key), (n, L[n - 1])
    for n in (28, 40, 42, 55, 59):
        assert L[n - 1] == "", n
    assert L[59] == "=" * 76 and L[1] == "=" * 76
    # the metadata values start in column 50
    assert L[30][:49].rstrip() == "Longitude (degrees):" and len(L[30]) == 49 + 11
    assert L[61].startswith(fullchem.species[0] + ",")
    assert L[61 + fullchem.nspec].startswith("R1,")
    assert L[61 + fullchem.nspec + fullchem.nreact].startswith("A1,")
    assert L[-1].startswith("A%d," % fullchem.nreact)


def _fortran(values, ints):
    if not FMT.exists():
        pytest.skip("oracle/_build/kppsa_fmt not built (make -C oracle fmt)")
    inp = "".join("R %d\n" % struct.unpack("<q", struct.pack("<d", float(v)))[0] for v in values)
    inp += "".join("I %d\n" % i for i in ints)
    out = subprocess.run([str(FMT)], input=inp, capture_output=True, text=True, check=True).stdout
    lines = out.split("\n")
    reals = [lines[6 * k:6 * k + 6] for k in range(len(values))]
    return reals, lines[6 * len(values):6 * len(values) + len(ints)]


FMT_VALUES = [0.0, 1.0, -1.0, 0.5, 123.456789, -0.00004999, 0.00005, 2.4e19, -3.3e-7, 1e100, 1e-100,
              -1e-300, 1.7976931348623157e308, 5e-324, 1e25, 1e-2, 9.99999999999999999e-3, 99999.99995,
              1234567.0, 1200.0, 0.1, 1.0 / 3.0, -2.0 / 3.0, 6.02214076e23, 288.155, 2.5e-299]


def test_numbers_as_the_fortran_runtime_writes_them(hip_lib, fullchem, tmp_path):
    """Every value through each field of the file against flang's rendering
    of the same edit descriptor."""
    hip = hip_lib
    reals, ints = _fortran(FMT_VALUES, [0, 1, -1, 17, 99999, 200000, -123456])
    nspec, nreact = fullchem.nspec, fullchem.nreact
    for k, v in enumerate(FMT_VALUES):
        es25, es10, f114, f112, e114, f136 = reals[k]
        conc = np.full(nspec, v)
        atol = np.full(nspec, v)
        rct = np.full(nreact, v)
        meta = _meta(hip, lon=v, lat=v, pressure_hpa=v, temperature=v, airnumden=v, h2o_vmr=v,
                     cloud_frac=v, suncos=v, h_init=v, h_exit=v, dt=v)
        for j in range(20):
            meta.rcntrl[j] = v
        p = tmp_path / ("v%d.txt" % k)
        hip.kppsa_write(fullchem, p, meta, conc, atol, rct, rct, names=fullchem.species)
        L = p.read_text().split("\n")
        val = lambda n: L[n - 1][49:]
        assert val(31) == f114 and val(32) == f114 and val(34) == f114, v
        assert val(35) == f112, v
        assert val(36) == e114 and val(37) == e114 and val(38) == e114 and val(39) == e114, v
        assert val(43) == f114 and val(44) == f114 and val(45) == f114, v
        assert L[50] == f136 * 5 and L[53] == f136 * 5, v
        name, a, b = L[61].split(",")
        assert (a, b) == (es25, es10), v
        assert L[61 + nspec] == "R1," + es25, v
    for k, i in enumerate([0, 1, -1, 17, 99999, 200000, -123456]):
        meta = _meta(hip, level=i, nsteps=i)
        for j in range(20):
            meta.icntrl[j] = i
        p = tmp_path / ("i%d.txt" % k)
        z = np.zeros(nspec)
        hip.kppsa_write(fullchem, p, meta, z, z, np.zeros(nreact), np.zeros(nreact))
        L = p.read_text().split("\n")
        assert L[32][49:] == ints[k] and L[45][49:] == ints[k], i
        assert L[47] == ints[k] * 10 and L[48] == ints[k] * 10, i
        assert L[61].startswith("C1,")


def test_read_rejects_a_file_of_another_mechanism(hip_lib, mech_of, tmp_path):
    hip = hip_lib
    hg = mech_of("Hg")
    conc, atol, rct, aout = _values(hg, 3)
    p = tmp_path / "hg.txt"
    hip.kppsa_write(hg, p, _meta(hip), conc, atol, rct, aout, names=hg.species)
    with pytest.raises(hip.GckppError, match="expected R1|data lines|species line"):
        hip.kppsa_read(mech_of("fullchem"), p)
    with pytest.raises(hip.GckppError, match="cannot open"):
        hip.kppsa_read(hg, tmp_path / "missing.txt")
