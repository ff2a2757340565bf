"""The oracle's restatement of fullchem_SetStateHet (oracle/setstatehet.c,
KPP/fullchem/fullchem_HetStateFuncs.F90:48-616), pinned on the CPU.

  * Cld_Params, SafeDiv and the Henry's-law ratio of Compute_L2G_Local
    against the reference's own routines (rateLawUtilFuncs.F90, henry_mod.F90)
    compiled by oracle/Makefile: bitwise, over every branch.
  * The rest -- Get_Theta_Ice, Halide_Conc, Get_Halide_CldConc,
    Get_Halide_SsaConc and the record itself -- by what the physics fixes:
    the in-cloud halide is the Henry's-law dissolved fraction of the total
    gas + aerosol halide (computed here from the compiled L2G), split in
    proportion to its sources; the sea-salt halide is moles per litre of
    aerosol water; Langmuir coverages add up with the free sites to one and
    follow van 't Hoff in the low-coverage limit; pH maps to [H+]; the
    alkalinity fraction is clipped to [0, 1]; the PSC flags follow their
    truth table.  A mistake planted in a copy of the record (a transposed
    source, a dropped 0.7) fails these checks.
"""
import numpy as np
import pytest

from gckpp_amd import fields, synth

AVO = 6.022140857e+23
F = fields.HET
G = fields.HIN


def _ref():
    from oracle import ref
    if not ref.available("fullchem"):
        pytest.skip("compiled reference not built")
    return ref


def _orc():
    from oracle import orc
    if not orc.lib_path().exists():
        pytest.skip("oracle not built")
    return orc


def test_cld_params_safediv_l2g_match_the_reference():
    ref, orc = _ref(), _orc()
    rng = np.random.default_rng(5)
    n = 400
    ad = 10 ** rng.uniform(6, 11, n)
    cldf = np.where(rng.uniform(0, 1, n) < 0.2, 0.0, rng.uniform(-0.1, 1.0, n))
    frl, fro = rng.uniform(0, 1, n), rng.uniform(0, 1, n)
    qi = np.where(rng.uniform(0, 1, n) < 0.3, 0.0, 10 ** rng.uniform(-8, -4, n))
    ql = np.where(rng.uniform(0, 1, n) < 0.3, 0.0, 10 ** rng.uniform(-8, -3, n))
    ql[:5] = 0.0
    qi[:5] = 0.0
    t = rng.uniform(180, 310, n)
    t[5:8] = (202.0, 217.0, 201.999)
    vair = 10 ** rng.uniform(14, 18, n)
    want = ref.cld_params(ad, cldf, frl, fro, qi, ql, t, vair)
    got = np.array([orc.cld_params(*a) for a in zip(ad, cldf, frl, fro, qi, ql, t, vair)]).T
    assert np.array_equal(got, want)
    # SafeDiv: ordinary quotients, a zero denominator, overflow and underflow
    # of the exponent difference, zero, Inf and NaN operands
    num = np.array([1.0, 1.0, 1e300, 1e-300, 0.0, np.inf, 3.0, 1e-310, np.nan, 2.5, -1e200])
    den = np.array([3.0, 0.0, 1e-300, 1e300, 2.0, 1.0, np.inf, 1e300, 1.0, np.nan, 1e-200])
    alt = np.full(num.size, 7.0)
    want = ref.safediv(num, den, alt)
    got = np.array([orc.safediv(a, b, c) for a, b, c in zip(num, den, alt)])
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.array_equal(got[ok], want[ok])
    # Henry's law with and without the pH correction
    tk = rng.uniform(200, 310, n)
    ph = np.where(rng.uniform(0, 1, n) < 0.2, rng.uniform(-2, 0, n), rng.uniform(0, 8, n))
    h2o = 10 ** rng.uniform(-12, -5, n)
    for k0, cr, pka in ((1.0, 9000.0, -6.3), (0.75, 10200.0, -9.0), (2.0, 500.0, -150.0)):
        want = ref.l2g(k0, cr, pka, tk, h2o, ph)
        got = np.array([orc.l2g(k0, cr, pka, a, b, c) for a, b, c in zip(tk, h2o, ph)])
        assert np.array_equal(got, want)


def _case(m, n=600, seed=11):
    rng = np.random.default_rng(seed)
    hin = synth.model_hetin(rng, n)
    conc = synth.box_cells(m, n, seed).conc.copy()
    sp = {s: m.species.index(s) for s in ("HBr", "BrSALA", "BrSALC", "HCl", "SALACL", "SALCCL",
                                         "NIT", "NITs", "SO4", "SO2", "HNO3")}
    # halide sources absent in some cells
    for s in ("HBr", "BrSALA", "BrSALC", "HCl", "SALACL", "SALCCL"):
        conc[sp[s], rng.uniform(0, 1, n) < 0.1] = 0.0
    conc[sp["HBr"], :8] = 0.0
    conc[sp["BrSALA"], :8] = 0.0
    conc[sp["BrSALC"], :8] = 0.0
    return hin, conc, sp


def _check_record(m, hin, conc, sp, het, naero=14):
    """The physical checks of one HetState record; returns nothing, asserts."""
    from oracle import ref
    I = lambda k, j=0: hin[G[k][0] + j]
    H = lambda k, j=0: het[F[k][0] + j]
    C = lambda s: conc[sp[s]]
    rtol = 4e-15
    close = lambda a, b: np.all(np.abs(a - b) <= rtol * np.maximum(np.abs(a), np.abs(b)) + 1e-300)
    # meteorology and cloud
    cf = np.clip(I("CLDF"), 0.0, 1.0)
    assert np.array_equal(H("cldFr"), cf) and np.array_equal(H("clearFr"), 1.0 - cf)
    assert np.array_equal(H("vAir"), I("AIRVOL") * 1e6)
    cld = ref.cld_params(I("AD"), I("CLDF"), I("FRLAND"), I("FROCEAN"), I("QI"), I("QL"),
                         I("T"), H("vAir"))
    for k, name in enumerate(("rLiq", "rIce", "aLiq", "vLiq", "aIce", "vIce")):
        assert np.array_equal(H(name), cld[k]), name
    # aerosol: areas, radii, volumes = A r / 3, water in the record's unit
    for j in range(14):
        if j < naero:
            assert np.array_equal(H("xArea", j), I("AeroArea", j))
            assert close(H("xVol", j), I("AeroArea", j) * I("AeroRadi", j) / 3.0)
            assert close(H("xH2O", j), I("AeroH2O", j) * 1e-6)
        else:
            assert (H("xArea", j) == 0).all() and (H("xVol", j) == 0).all()
    # pH -> [H+]
    with np.errstate(divide="ignore"):
        assert np.allclose(-np.log10(H("H_conc_Sul")), I("IsorropAeropH"), rtol=0, atol=1e-13)
        assert np.allclose(-np.log10(H("H_conc_LCl")), I("pHCloud"), rtol=0, atol=1e-13)
    assert abs(-np.log10(H("H_conc_ICl")[0]) - 4.5) < 1e-14
    assert (H("H_conc_SSC") == 1e-5).all() and np.array_equal(H("H_conc_SSA"), H("H_conc_Sul"))
    # alkalinity: the clipped ratio, flags complementary, threshold 1 %
    for ss, al in (("SSA", ("SALAAL", "SALA")), ("SSC", ("SALCAL", "SALC"))):
        ratio = np.where(I(al[1]) > 0, I(al[0]) / np.where(I(al[1]) > 0, I(al[1]), 1.0), 0.0)
        f = np.clip(ratio, 0.0, 1.0)
        assert close(H("f_Alk_" + ss), f)
        assert np.array_equal(H("f_Acid_" + ss), 1.0 - H("f_Alk_" + ss))
        assert np.array_equal(H(ss + "_is_Alk"), (H("f_Alk_" + ss) > 0.01).astype(float))
        assert np.array_equal(H(ss + "_is_Acid") + H(ss + "_is_Alk"), np.ones(het.shape[1]))
    # S(IV): fractions of the total; HSO3- in molec per cm3 of air (mol/L x
    # 1e-3 L/cm3 x AVO x cloud water QL*AIRDEN*1e-3) in the cloudy part, per
    # molecule of SO2
    tso3 = I("HSO3_aq") + I("SO3_aq")
    pos = tso3 > 0
    assert close(H("frac_HSO3_aq")[pos] + H("frac_SO3_aq")[pos], 1.0)
    assert (H("frac_HSO3_aq")[~pos] == 0).all()
    ok = (I("CLDF") > 0) & (C("SO2") > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        want_m = I("HSO3_aq") * 1e-3 * I("QL") * I("AIRDEN") * 1e-3 * AVO / I("CLDF") / C("SO2")
    assert np.allclose(H("HSO3m")[ok], want_m[ok], rtol=1e-14, atol=0)
    assert (H("HSO3m")[I("CLDF") == 0] == 0).all()
    # Langmuir coverages of HNO3 and HCl on ice: with the free sites they sum
    # to one, and in the low-coverage limit ln(theta/c) follows van 't Hoff
    # with 4585 K and 4600 K
    free = 1.0 - H("HNO3_theta") - H("HCl_theta")
    assert (free > 0).all() and (H("HNO3_theta") >= 0).all()
    t = I("T")
    k1 = H("HNO3_theta") / free / np.where(C("HNO3") > 0, C("HNO3"), 1.0)
    k2 = H("HCl_theta") / free / np.where(C("HCl") > 0, C("HCl"), 1.0)
    sel = (C("HNO3") > 0) & (C("HCl") > 0)
    slope1 = np.polyfit(1.0 / t[sel], np.log(k1[sel]), 1)[0]
    slope2 = np.polyfit(1.0 / t[sel], np.log(k2[sel]), 1)[0]
    assert abs(slope1 - 4585.0) < 1e-6 and abs(slope2 - 4600.0) < 1e-6
    # HBr on ice: a power law c^0.88 capped at full coverage
    hb = C("HBr")
    s = (hb > 1e3) & (H("HBr_theta") < 1.0)
    sl = np.polyfit(np.log(hb[s]), np.log(H("HBr_theta")[s]), 1)[0]
    assert abs(sl - 0.88) < 1e-9 and (H("HBr_theta") <= 1.0).all()
    # in-cloud halides: the dissolved fraction L2G/(1+L2G) of the total
    # (gas + 0.7 fine + coarse) halide, per litre of cloud water, split in
    # proportion to the three sources
    vtot = np.where(cf > 0, H("vLiq") / np.where(cf > 0, cf, 1.0), 0.0)
    wet = vtot >= 1e-20
    for X, srcs, (k0, cr, pka) in (("Br", ("HBr", "BrSALA", "BrSALC"), (0.75, 10200.0, -9.0)),
                                   ("Cl", ("HCl", "SALACL", "SALCCL"), (1.0, 9000.0, -6.3))):
        g, a, c = (C(s_) for s_ in srcs)
        tot = g + 0.7 * a + c
        l2g = ref.l2g(k0, cr, pka, t, np.where(wet, vtot, 1.0), I("pHCloud"))
        conc_x = np.where(wet, l2g / (1.0 + l2g) * tot / (np.where(wet, vtot, 1.0) * AVO * 1e-3), 0.0)
        conc_x = np.where(tot > 0, conc_x, 0.0)
        parts = H(X + "_conc_CldG"), H(X + "_conc_CldA"), H(X + "_conc_CldC")
        assert np.allclose(H(X + "_conc_Cld"), conc_x, rtol=1e-13, atol=0), X
        q = tot > 0
        for p_, src in zip(parts, (g, 0.7 * a, c)):
            assert np.allclose(p_[q], (conc_x * src / np.where(q, tot, 1.0))[q], rtol=1e-13, atol=0)
        fr = [H("frac_%s_Cld%s" % (X, z)) for z in "GAC"]
        pos = H(X + "_conc_Cld") > 0
        assert close(sum(fr)[pos], 1.0) and all((f_[~pos] == 0).all() for f_ in fr)
    # sea-salt halides and nitrate: moles per litre of aerosol water
    for name, src, area, radi in (("Br_conc_SSA", "BrSALA", I("aClArea"), I("aClRadi")),
                                  ("Cl_conc_SSA", "SALACL", I("aClArea"), I("aClRadi")),
                                  ("NIT_conc_SSA", "NIT", I("aClArea"), I("aClRadi")),
                                  ("Br_conc_SSC", "BrSALC", H("xArea", 11), H("xRadi", 11)),
                                  ("Cl_conc_SSC", "SALCCL", H("xArea", 11), H("xRadi", 11)),
                                  ("NIT_conc_SSC", "NITs", H("xArea", 11), H("xRadi", 11))):
        litres = area * radi / 3.0 * 1e-3
        want = np.where(litres > 1e-20, np.maximum(C(src) / AVO / np.where(litres > 0, litres, 1.0), 0), 0)
        assert np.allclose(H(name), want, rtol=1e-14, atol=0), name
    for z, num, den in (("Cld", "Br_conc_Cld", "Cl_conc_Cld"), ("SSA", "Br_conc_SSA", "Cl_conc_SSA"),
                        ("SSC", "Br_conc_SSC", "Cl_conc_SSC")):
        q = H(den) > 0
        assert close(H("Br_over_Cl_" + z)[q], H(num)[q] / H(den)[q]) and (H("Br_over_Cl_" + z)[~q] == 0).all()
    tot = C("SALACL") + C("NIT") + C("SO4")
    assert close(H("frac_SALACL"), np.where(tot > 0, C("SALACL") / np.where(tot > 0, tot, 1), 0))
    # stratosphere and PSC flags
    strat = I("InStratosphere") != 0
    psc = strat & (I("STATE_PSC") >= 2)
    assert np.array_equal(H("stratBox"), strat.astype(float))
    assert np.array_equal(H("pscBox"), psc.astype(float))
    assert np.array_equal(H("natSurface"), (psc & (C("NIT") > 0)).astype(float))
    for j in range(11):
        assert np.array_equal(H("KHETI_SLA", j), I("KHETI_SLA", j))


def test_set_state_het_physics(fullchem):
    _ref(), _orc()
    from oracle import orc
    hin, conc, sp = _case(fullchem)
    het = orc.set_state_het(fullchem, hin, conc)
    assert not np.isnan(het).any()
    _check_record(fullchem, hin, conc, sp, het)
    # nAeroType below 12: no coarse sea salt, its halides zero
    het8 = orc.set_state_het(fullchem, hin, conc, naerotype=8)
    _check_record(fullchem, hin, conc, sp, het8, naero=8)
    assert (het8[F["Br_conc_SSC"][0]] == 0).all()


@pytest.mark.parametrize("mistake", ["swap_sources", "drop_07", "theta_sign"])
def test_checks_catch_a_planted_mistake(fullchem, mistake):
    """A copy of the oracle's record with one plausible mistake fails them."""
    _ref(), _orc()
    from oracle import orc
    hin, conc, sp = _case(fullchem, n=300, seed=3)
    het = orc.set_state_het(fullchem, hin, conc)
    bad = het.copy()
    if mistake == "swap_sources":
        a, c = F["Br_conc_CldA"][0], F["Br_conc_CldC"][0]
        bad[[a, c]] = bad[[c, a]]
    elif mistake == "drop_07":
        conc2 = conc.copy()
        conc2[sp["SALACL"]] /= 0.7          # as if the 0.7 were missing
        bad = orc.set_state_het(fullchem, hin, conc2)
    else:
        bad[F["HCl_theta"][0]] = -bad[F["HCl_theta"][0]] + 2 * bad[F["HCl_theta"][0]] ** 2
    with pytest.raises(AssertionError):
        _check_record(fullchem, hin, conc, sp, bad)
