"""The oracle's restatement of fullchem_SulfurAqChem (oracle/producers.c,
KPP/fullchem/fullchem_SulfurChemFuncs.F90:120-305), pinned on the CPU.

Its building blocks, Ars_L1k and kIIR1Ltd, against the reference's own
routines compiled by oracle/Makefile (rateLawUtilFuncs.F90:77-150), bitwise,
over every branch (no aerosol, educt below one molecule, the lifetime
limits); then K_MT(1:6) against those compiled blocks composed as the
subroutine's six branches say: fine sea salt (type 11) for K_MT(1:3),
coarse (type 12) for K_MT(4:6), SO2 only where O3 > 1e10 and per O3
molecule, nothing where the alkalinity is at most 0.1.  A mistake planted
in the oracle's output (fine and coarse swapped) fails the check.
"""
import numpy as np
import pytest

from gckpp_amd import fields, synth


def _libs():
    from oracle import orc, ref
    if not ref.available("fullchem") or not orc.lib_path().exists():
        pytest.skip("oracles not built")
    return orc, ref


def test_uptake_blocks_match_reference():
    orc, ref = _libs()
    rng = np.random.default_rng(2)
    n = 500
    area = np.where(rng.uniform(0, 1, n) < 0.1, 0.0, 10 ** rng.uniform(-9, -4, n))
    radius = np.where(rng.uniform(0, 1, n) < 0.05, 0.0, 10 ** rng.uniform(-6, -3, n))
    gamma = np.where(rng.uniform(0, 1, n) < 0.05, 0.0, rng.uniform(0.01, 1.0, n))
    srmw = np.sqrt(rng.uniform(10, 150, n))
    numden = 10 ** rng.uniform(17, 19.5, n)
    temp = rng.uniform(180, 310, n)
    cgas = 10 ** rng.uniform(-2, 12, n)
    ceduct = np.where(rng.uniform(0, 1, n) < 0.1, rng.uniform(0, 1, n), 10 ** rng.uniform(-1, 14, n))
    want = ref.uptake(area, radius, gamma, srmw, numden, temp, cgas, ceduct)
    L = orc._lib()
    L.orc_ars_l1k.argtypes = [orc.ctypes.c_double] * 6
    L.orc_ars_l1k.restype = orc.ctypes.c_double
    L.orc_kiir1ltd.argtypes = [orc.ctypes.c_double] * 3
    L.orc_kiir1ltd.restype = orc.ctypes.c_double
    k1 = np.array([L.orc_ars_l1k(a, r, g, s, d, np.sqrt(t))
                   for a, r, g, s, d, t in zip(area, radius, gamma, srmw, numden, temp)])
    k2 = np.array([L.orc_kiir1ltd(g, e, k) for g, e, k in zip(cgas, ceduct, k1)])
    assert np.array_equal(k1, want[0]) and np.array_equal(k2, want[1])
    # every branch of kIIR1Ltd is reached
    assert (want[1] == 0).any() and (want[1] > 0).any()


def _case(m, n, seed):
    rng = np.random.default_rng(seed)
    w = synth.box_cells(m, n, seed)
    hin = synth.model_hetin(rng, n)
    conc = w.conc.copy()
    ix = m.species.index
    conc[ix("SALAAL")] = np.where(rng.uniform(0, 1, n) < 0.3, rng.uniform(0, 0.2, n),
                                  10 ** rng.uniform(-1, 8, n))
    conc[ix("SALCAL")] = np.where(rng.uniform(0, 1, n) < 0.3, rng.uniform(0, 0.2, n),
                                  10 ** rng.uniform(-1, 8, n))
    conc[ix("O3")] = 10 ** rng.uniform(9, 13, n)
    conc[ix("O3"), :4] = 1.0e10
    conc[ix("SALAAL"), 4:8] = 0.1
    return w, hin, conc


def _expected(ref, m, w, hin, conc):
    ix = m.species.index
    mw = np.asarray(m.species_consts["mw"], dtype=np.float64)
    srmw = lambda s: np.sqrt(mw[ix(s)])
    T, ND = w.met[fields.MET["TEMP"][0]], w.met[fields.MET["NUMDEN"][0]]
    o3 = conc[ix("O3")]
    out = np.zeros((6, conc.shape[1]))
    for f, alk in ((0, "SALAAL"), (1, "SALCAL")):
        area = hin[fields.HIN["WetAeroArea"][0] + 10 + f]
        radi = hin[fields.HIN["AeroRadi"][0] + 10 + f]
        on = conc[ix(alk)] > 0.1
        for j, (gas, gam) in enumerate((("SO2", 0.11), ("HCl", 0.07), ("HNO3", 0.5))):
            k = ref.uptake(area, radi, gam, srmw(gas), ND, T, conc[ix(gas)], conc[ix(alk)])[1]
            if gas == "SO2":
                out[3 * f + j] = np.where(on & (o3 > 1e10), k / o3, 0.0)
            else:
                out[3 * f + j] = np.where(on, k, 0.0)
    return out


def _check(ref, m, w, hin, conc, kmt):
    want = _expected(ref, m, w, hin, conc)
    assert np.array_equal(kmt, want)


def test_sulfur_aq_matches_composed_reference_blocks(fullchem):
    orc, ref = _libs()
    w, hin, conc = _case(fullchem, 800, 4)
    met = orc.sulfur_aq(fullchem, hin, conc, w.met)
    o = fields.MET["K_MT"][0]
    kmt = met[o:o + 6]
    _check(ref, fullchem, w, hin, conc, kmt)
    assert (kmt > 0).any(axis=1).all() and (kmt == 0).any(axis=1).all()
    # the other rows of met untouched
    rest = np.ones(fields.NMET, dtype=bool)
    rest[o:o + 6] = False
    assert np.array_equal(met[rest], w.met[rest])


def test_planted_mistake_is_caught(fullchem):
    orc, ref = _libs()
    w, hin, conc = _case(fullchem, 300, 6)
    met = orc.sulfur_aq(fullchem, hin, conc, w.met)
    o = fields.MET["K_MT"][0]
    bad = met[o:o + 6][[3, 4, 5, 0, 1, 2]]
    with pytest.raises(AssertionError):
        _check(ref, fullchem, w, hin, conc, bad)
