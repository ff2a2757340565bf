"""fullchem_SetStateHet and fullchem_SulfurAqChem on the GPU
(gckpp_set_state_het_dev, gckpp_sulfur_aq_dev) against the oracle's
restatements (oracle/producers.c, pinned by tests/test_oracle_setstatehet.py
and tests/test_oracle_sulfur_aq.py), and the chain SulfurAqChem +
SetStateHet -> Update_RCONST against the compiled reference's rates on the
oracle's records.

Tolerance: pass-through fields, flags and exact zeros bitwise; the computed
fields (exp, pow, sqrt and a few divisions per field, the device's libm
against the host's) to 1e-14 relative.  Rates: the bound of the rich-HetState
rate test, 1e-7 (the cloud-uptake laws cancel nearly equal terms), plus
1e-7 of each reaction's largest rate for the few tiny rates where that
cancellation meets fused multiply-adds.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _inputs(m, n, seed):
    from gckpp_amd import fields, synth
    rng = np.random.default_rng(seed)
    w = synth.box_cells(m, n, seed)
    hin = synth.model_hetin(rng, n)
    hin[fields.HIN["T"][0]] = w.met[fields.MET["TEMP"][0]]
    conc = w.conc.copy()
    for s in ("HBr", "BrSALA", "BrSALC", "HCl", "SALACL", "SALCCL", "NIT"):
        conc[m.species.index(s), rng.uniform(0, 1, n) < 0.1] = 0.0
    return w, hin, conc


@pytest.mark.parametrize("naero", [14, 8])
def test_set_state_het_matches_oracle(fullchem, T, naero):
    print("case: SetStateHet GPU vs oracle, nAeroType %d" % naero, flush=True)
    from gckpp_amd import fields, hip
    from oracle import orc
    n = 5000
    w, hin, conc = _inputs(fullchem, n, 21)
    want = orc.set_state_het(fullchem, hin, conc, naerotype=naero, gamma_ho2=0.2)
    ctx = hip.Context(fullchem, n)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    got = ctx.set_state_het(dev(hin), dev(conc), naerotype=naero, gamma_ho2=0.2).cpu().numpy()
    ctx.close()
    assert np.isfinite(got).all()
    assert np.array_equal(got == 0.0, want == 0.0)
    nz = want != 0.0
    rel = np.zeros_like(want)
    rel[nz] = np.abs(got[nz] - want[nz]) / np.abs(want[nz])
    names = {o + j: k for k, (o, l) in fields.HET.items() for j in range(l)}
    worst = np.unravel_index(rel.argmax(), rel.shape)
    print("max rel %.3g in %s" % (rel.max(), names[worst[0]]), flush=True)
    assert rel.max() <= 1e-14
    exact = [k for k in fields.HET if k.endswith(("Box", "_is_Alk", "_is_Acid", "Surface"))] + \
        ["xArea", "xRadi", "wetArea", "KHETI_SLA", "cldFr", "clearFr", "vAir", "NAEROTYPE",
         "pHCloud", "pHSSA", "HSO3_aq", "SO3_aq", "TSO3_aq", "rLiq"]
    for k in exact:
        o, l = fields.HET[k]
        assert np.array_equal(got[o:o + l], want[o:o + l]), k


def _alkalinity(m, conc, seed):
    rng = np.random.default_rng(seed)
    n = conc.shape[1]
    for s in ("SALAAL", "SALCAL"):
        conc[m.species.index(s)] = np.where(rng.uniform(0, 1, n) < 0.3, rng.uniform(0, 0.2, n),
                                            10 ** rng.uniform(-1, 8, n))
    conc[m.species.index("O3"), :64] = 1.0e10


def test_sulfur_aq_matches_oracle(fullchem, T):
    print("case: SulfurAqChem K_MT GPU vs oracle", flush=True)
    from gckpp_amd import fields, hip
    from oracle import orc
    n = 5000
    w, hin, conc = _inputs(fullchem, n, 41)
    _alkalinity(fullchem, conc, 41)
    want = orc.sulfur_aq(fullchem, hin, conc, w.met)
    ctx = hip.Context(fullchem, n)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    met_d = dev(w.met)
    got = ctx.sulfur_aq(dev(hin), dev(conc), met_d).cpu().numpy()
    ctx.close()
    o = fields.MET["K_MT"][0]
    rest = np.ones(fields.NMET, dtype=bool)
    rest[o:o + 6] = False
    assert np.array_equal(got[rest], w.met[rest])
    g, e = got[o:o + 6], want[o:o + 6]
    assert np.array_equal(g == 0, e == 0) and (e > 0).any(axis=1).all()
    nz = e != 0
    rel = np.abs(g[nz] - e[nz]) / np.abs(e[nz])
    print("K_MT max rel %.3g" % rel.max(), flush=True)
    assert rel.max() <= 1e-14


def test_set_state_het_then_rates_match_reference(fullchem, T):
    """The records the GPU producers write, fed to the GPU rates and to the
    compiled reference's Update_RCONST.  Bound per rate: 1e-7 relative (the
    rich-HetState rate test's) plus 1e-7 of the reaction's largest rate in
    the batch.  The second term covers CloudHet's entrainment limit, whose
    formula subtracts sqrt(...) ~ kk from kk for strongly taken-up gases
    (kk = 3600 kI): there the rounding of fused multiply-adds (the device
    contracts a*b+c within an expression, the reference's x86 build does
    not) moves the result by up to ~1e-4 relative, on rates that are at most
    2.5e-8 of the reaction's largest (seen: HOBr uptake by HCl/HBr, ClNO3 by
    HBr; the same rate-law source built for the host without contraction
    gives the reference's values bit for bit, with contraction the GPU's)."""
    print("case: SulfurAqChem + SetStateHet -> Update_RCONST on the GPU vs compiled reference",
          flush=True)
    from gckpp_amd import hip
    from oracle import ref
    if not ref.available("fullchem"):
        pytest.skip("compiled reference not built")
    n = 2048
    w, hin, conc = _inputs(fullchem, n, 33)
    _alkalinity(fullchem, conc, 33)
    ctx = hip.Context(fullchem, n)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    met_d = dev(w.met)
    ctx.sulfur_aq(dev(hin), dev(conc), met_d)
    het_d = ctx.set_state_het(dev(hin), dev(conc))
    got = ctx.rates(met_d, dev(w.photol), het_d, dev(conc)).cpu().numpy()
    ctx.close()
    w.met, w.het, w.conc = met_d.cpu().numpy(), het_d.cpu().numpy(), conc
    want = ref.RefKPP(fullchem).rates(w)
    assert np.isfinite(got).all()
    assert np.array_equal(got == 0.0, want == 0.0)
    err = np.abs(got - want)
    nz = want != 0.0
    rel = err[nz] / np.abs(want[nz])
    scale = np.abs(want).max(axis=1, keepdims=True)
    print("rates max rel %.3g, median rel %.3g, entries beyond 1e-7 relative %d of %d, "
          "largest of those against the reaction's scale %.3g"
          % (rel.max(), np.median(rel), int((rel > 1e-7).sum()), rel.size,
             (err / scale)[np.abs(got - want) > 1e-7 * np.abs(want)].max(initial=0.0)), flush=True)
    assert np.median(rel) <= 1e-15
    assert (err <= 1e-7 * np.abs(want) + 1e-7 * scale).all()


def test_host_pointer_producers_equal_device(fullchem, T):
    """gckpp_set_state_het / gckpp_sulfur_aq on host arrays, 2,500 cells in
    chunks of 1,000 staged through the device: bitwise the device calls on
    the whole batch; the host met keeps every row but K_MT."""
    print("case: host-pointer SetStateHet and SulfurAqChem, chunked", flush=True)
    from gckpp_amd import fields, hip
    n = 2500
    w, hin, conc = _inputs(fullchem, n, 51)
    _alkalinity(fullchem, conc, 51)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    big = hip.Context(fullchem, n)
    het_d = big.set_state_het(dev(hin), dev(conc), naerotype=12, gamma_ho2=0.1).cpu().numpy()
    met_d = big.sulfur_aq(dev(hin), dev(conc), dev(w.met)).cpu().numpy()
    big.close()
    small = hip.Context(fullchem, 1000)
    het_h = small.set_state_het_host(hin, conc, naerotype=12, gamma_ho2=0.1)
    met_h = small.sulfur_aq_host(hin, conc, w.met)
    small.close()
    assert np.array_equal(het_h, het_d)
    assert np.array_equal(met_h, met_d)
    o = fields.MET["K_MT"][0]
    rest = np.ones(fields.NMET, dtype=bool)
    rest[o:o + 6] = False
    assert np.array_equal(met_h[rest], w.met[rest])


def test_set_state_het_rejects_bad_arguments(fullchem, hg, T):
    from gckpp_amd import fields, hip
    n = 64
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = hip.Context(fullchem, n)
    hin = dev(np.zeros((fields.NHIN, n)))
    conc = dev(np.zeros((fullchem.nspec, n)))
    with pytest.raises(hip.GckppError, match="nAeroType"):
        ctx.set_state_het(hin, conc, naerotype=15)
    ctx.close()
    ctx = hip.Context(hg, n)
    with pytest.raises(hip.GckppError, match="fullchem"):
        ctx.set_state_het(hin, dev(np.zeros((hg.nspec, n))))
    with pytest.raises(hip.GckppError, match="fullchem"):
        ctx.sulfur_aq(hin, dev(np.zeros((hg.nspec, n))), dev(np.zeros((fields.NMET, n))))
    with pytest.raises(hip.GckppError, match="fullchem"):
        ctx.set_species_map(np.full(hg.nspec, 30.0), np.arange(hg.nspec), alk_equiv=True)
    ctx.close()
