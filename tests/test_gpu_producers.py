"""Do_FullChem's model side on the GPU: the gather of the KPP vector from the
model's species in kg/kg dry air with Convert_Spc_Units' conversion to
molec/cm3 (GeosUtil/unitconv_mod.F90:1287-1391, GeosCore/fullchem_mod.F90:
817-821), the prod/loss dummies set to zero (:942-946), and the scatter back
with MAX(C,0) (:1329-1350) and the conversion to kg/kg dry (:1407-1515).

Checks: the gather and scatter against the same products evaluated in numpy
in the reference's order (bitwise); air itself (MW 28.9644 g/mol, 1 kg/kg)
gathered at a density from the ideal gas law comes out as the number
density P N_A/(R T) (to 1e-14); gather then scatter returns the model
species (clipped) to 8e-16 relative (seven roundings: two each way and
the factors' product, (AVO/MW)/1e6 times 1e6/(AVO/MW), within three of
one) and leaves the species KPP does not hold untouched; a chemistry step between them equals the step on the
numpy-gathered vector.  With the alkalinity conversion
(fullchem_ConvertAlkToEquiv / fullchem_ConvertEquivToAlk,
KPP/fullchem/fullchem_SulfurChemFuncs.F90:67-105) SALAAL and SALCAL are
gathered times MW * 7e-5 and scattered back divided by it."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

AVO = 6.022140857e+23          # Headers/physconstants.F90:32
BOLTZ = 1.38064852e-23         # :67
RSTARG = 8.3144598             # :72
AIRMW = 28.9644                # :25


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _model(m, n, seed):
    """A model state: every KPP species among nmodel = NSPEC + 6 model
    species in a shuffled order (the last model species is air), two KPP
    species without a model species, kg/kg dry values, air densities."""
    rng = np.random.default_rng(seed)
    nmodel = m.nspec + 6
    slots = rng.permutation(nmodel - 1)
    kmap = slots[:m.nspec].astype(np.int32)
    kmap[[5, 17]] = -1
    mw = np.full(nmodel, 50.0)
    mw_kpp = np.asarray(m.species_consts["mw"], dtype=np.float64)
    for k in range(m.nspec):
        if kmap[k] >= 0:
            mw[kmap[k]] = mw_kpp[k] if mw_kpp[k] > 0 else 30.0
    mw[nmodel - 1] = AIRMW
    spc = 10.0 ** rng.uniform(-14, -5, (nmodel, n))
    spc[:, ::7] *= -1e-3                       # some negatives, clipped on the way back
    spc[nmodel - 1] = 1.0
    temp = rng.uniform(190, 310, n)
    press = 10.0 ** rng.uniform(2, 5, n)       # Pa
    airden = press * AIRMW * 1e-3 / (RSTARG * temp)
    zero = np.zeros(m.nspec, dtype=np.int32)
    zero[m.family_species()] = 1
    return kmap, mw, spc, airden, zero, press, temp


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
def test_gather_scatter_and_step(name, mech_of, T):
    print("case: gather/scatter with unit conversion, %s" % name, flush=True)
    from gckpp_amd import hip, options, synth
    m = mech_of(name)
    n = 3000
    kmap, mw, spc, airden, zero, press, temp = _model(m, n, 7)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = hip.Context(m, n)
    ctx.set_species_map(mw, kmap, zero)
    sp_d, ad_d = dev(spc), dev(airden)
    conc = ctx.gather(sp_d, ad_d).cpu().numpy()
    # numpy, the reference's order: Conc * AIRDEN * ((AVO / MW_kg) / 1e6)
    want = np.zeros((m.nspec, n))
    for k in range(m.nspec):
        s = kmap[k]
        if s >= 0 and not zero[k]:
            want[k] = spc[s] * airden * ((AVO / (mw[s] * 1e-3)) / 1e6)
    assert np.array_equal(conc, want)
    assert (conc[[5, 17]] == 0).all() and (conc[zero == 1] == 0).all()
    # air: the ideal-gas number density P N_A / (R T), which is P / (k_B T)
    # to the 7e-9 by which physconstants' k_B N_A and R differ
    ctx_air = hip.Context(m, n)
    amap = np.full(m.nspec, -1, dtype=np.int32)
    amap[0] = spc.shape[0] - 1
    ctx_air.set_species_map(mw, amap)
    nd = ctx_air.gather(sp_d, ad_d).cpu().numpy()[0]
    assert np.abs(nd / (press * AVO / (RSTARG * temp) * 1e-6) - 1.0).max() < 1e-14
    assert np.abs(nd / (press / (BOLTZ * temp) * 1e-6) - 1.0).max() < 1e-8
    ctx_air.close()
    # gather then scatter: the model species back (clipped), the others untouched
    back = sp_d.clone()
    ctx.scatter(dev(conc), ad_d, back)
    back = back.cpu().numpy()
    mapped = np.zeros(spc.shape[0], dtype=bool)
    for k in range(m.nspec):
        if kmap[k] >= 0 and not zero[k]:
            mapped[kmap[k]] = True
            exp = np.maximum(spc[kmap[k]], 0.0)
            assert np.all(np.abs(back[kmap[k]] - exp) <= 8.0e-16 * np.abs(exp)), k
    untouched = np.ones(spc.shape[0], dtype=bool)
    untouched[kmap[kmap >= 0]] = False
    assert np.array_equal(back[untouched], spc[untouched])
    # a chemistry step between gather and scatter: the same as on the
    # numpy-gathered vector with the numpy scatter
    w = synth.box_cells(m, n, 3)
    ic, rc = options.icntrl_for(name)
    c1 = ctx.gather(sp_d, ad_d)
    c2 = dev(want)
    for c in (c1, c2):
        ctx.fullchem_step(dev(w.met), dev(w.photol), dev(w.het), c, T.zeros(n, dtype=T.float64,
                          device="cuda"), w.dt, ic, rc)
    T.cuda.synchronize()
    assert T.equal(c1, c2)
    out = sp_d.clone()
    ctx.scatter(c1, ad_d, out)
    out = out.cpu().numpy()
    c2h = c2.cpu().numpy()
    for k in range(m.nspec):
        s = kmap[k]
        if s >= 0:
            exp = np.maximum(c2h[k], 0.0) * (1e6 / (AVO / (mw[s] * 1e-3))) / airden
            assert np.array_equal(out[s], exp), k
    ctx.close()


def test_alkalinity_equivalents(fullchem, T):
    print("case: gather/scatter with the alkalinity conversion, fullchem", flush=True)
    from gckpp_amd import hip
    m = fullchem
    n = 1000
    kmap, mw, spc, airden, zero, press, temp = _model(m, n, 9)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = hip.Context(m, n)
    ctx.set_species_map(mw, kmap, zero)
    sp_d, ad_d = dev(spc), dev(airden)
    plain = ctx.gather(sp_d, ad_d).cpu().numpy()
    ctx.set_species_map(mw, kmap, zero, alk_equiv=True)
    conc = ctx.gather(sp_d, ad_d).cpu().numpy()
    alk = [m.species.index("SALAAL"), m.species.index("SALCAL")]
    other = np.ones(m.nspec, dtype=bool)
    other[alk] = False
    assert np.array_equal(conc[other], plain[other])
    for k in alk:
        assert kmap[k] >= 0
        assert np.array_equal(conc[k], plain[k] * (mw[kmap[k]] * 7.0e-5))
    back = sp_d.clone()
    ctx.scatter(dev(conc), ad_d, back)
    back = back.cpu().numpy()
    for k in alk:
        s = kmap[k]
        exp = np.maximum(conc[k] / (mw[s] * 7.0e-5), 0.0) * (1e6 / (AVO / (mw[s] * 1e-3))) / airden
        assert np.array_equal(back[s], exp)
        assert np.all(np.abs(back[s] - np.maximum(spc[s], 0)) <= 1.0e-15 * np.abs(spc[s]))
    ctx.close()
