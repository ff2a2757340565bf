"""KPP-Standalone state files around a GPU step: the cells' state before
Integrate (concentrations, the GPU's rate constants, Fun's reaction rates
A from gckpp_aout_dev, the Hstart and options, the step's Hexit and step
count) written with gckpp_kppsa_write, read back with gckpp_kppsa_read, and
integrated again on the GPU from the files alone -- bitwise the first
result -- and by the compiled reference (oracle/_ref) from the same files,
to the repo tolerance.  The KPP Standalone replays a cell this way
(GeosCore/kppsa_interface_mod.F90:569-843 writes the file at
GeosCore/fullchem_mod.F90:1306-1321).  A is checked against mass action
(RCT times the reactant concentrations) evaluated in numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _mass_action(m, conc, rct):
    A = rct.copy()
    for r, rx in enumerate(m.reactions):
        for s, k in rx["reactants"]:
            for _ in range(int(k)):
                A[r] = A[r] * conc[s]
    return A


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
def test_state_files_replay_the_step(name, mech_of, T, tmp_path):
    print("case: KPP-Standalone files, %s" % name, flush=True)
    from gckpp_amd import hip, options, synth
    from gckpp_amd.fields import MET
    from oracle import ref
    m = mech_of(name)
    w = synth.box_cells(m, 4096, 0).subset(np.r_[0:16, 2048:2064])
    n = w.ncell
    w.hstart = np.where(np.arange(n) % 2 == 0, 0.0, 60.0 + 10.0 * np.arange(n))
    ic, rc = options.icntrl_for(name)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = hip.Context(m, n)
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    conc0 = dev(w.conc)
    rct = ctx.rates(dev(w.met), dev(w.photol), dev(w.het), conc0)
    aout = ctx.aout(conc0, rct)
    conc1 = conc0.clone()
    ist, rst, ierr = ctx.integrate(conc1, rct, dev(w.hstart), 0.0, w.dt, ic, rc)
    T.cuda.synchronize()
    c0, r0, a0 = conc0.cpu().numpy(), rct.cpu().numpy(), aout.cpu().numpy()
    c1, ist1, rst1, e1 = conc1.cpu().numpy(), ist.cpu().numpy(), rst.cpu().numpy(), ierr.cpu().numpy()
    assert (e1 == 1).all()
    ma = _mass_action(m, c0, r0)
    assert np.all(np.abs(a0 - ma) <= 1e-15 * np.abs(ma)), "Aout differs from mass action"

    atol = np.asarray(m.species_consts["atol"], dtype=np.float64)
    paths = []
    for i in range(n):
        meta = hip.KppsaMeta()
        meta.location = b"box_cell_%d" % i
        meta.timestamp = b"2019/06/21 12:00"
        meta.level = 1 + i % 72
        meta.temperature, meta.pressure_hpa = w.met[MET["TEMP"][0], i], w.met[MET["PRESS"][0], i]
        meta.airnumden, meta.suncos = w.met[MET["NUMDEN"][0], i], w.met[MET["SUNCOS"][0], i]
        meta.h2o_vmr = w.met[MET["H2O"][0], i] / w.met[MET["NUMDEN"][0], i]
        meta.h_init, meta.h_exit, meta.dt = w.hstart[i], rst1[1, i], w.dt
        meta.nsteps = int(ist1[2, i])
        for k in range(20):
            meta.icntrl[k] = int(ic[k])
            meta.rcntrl[k] = float(rc[k])
        meta.rcntrl[2] = w.hstart[i]
        p = tmp_path / ("cell%d.txt" % i)
        hip.kppsa_write(m, p, meta, c0[:, i], atol, r0[:, i], a0[:, i], names=m.species)
        paths.append(p)

    rd = [hip.kppsa_read(m, p) for p in paths]
    cf = np.stack([x[2] for x in rd], 1)
    rf = np.stack([x[4] for x in rd], 1)
    af = np.stack([x[5] for x in rd], 1)
    hf = np.array([x[0].h_init for x in rd])
    icf = np.array(list(rd[0][0].icntrl), dtype=np.int32)
    rcf = np.array(list(rd[0][0].rcntrl))
    rcf[2] = 0.0                            # per cell through hstart
    assert np.array_equal(cf, c0) and np.array_equal(rf, r0) and np.array_equal(af, a0)
    assert np.array_equal(hf, w.hstart) and np.array_equal(icf, np.asarray(ic, dtype=np.int32))
    assert [x[1] for x in rd][0] == list(m.species)
    assert [x[0].nsteps for x in rd] == [int(v) for v in ist1[2]]

    conc2 = dev(cf)
    ist2, rst2, e2 = ctx.integrate(conc2, dev(rf), dev(hf), 0.0, rd[0][0].dt, icf, rcf)
    T.cuda.synchronize()
    assert np.array_equal(conc2.cpu().numpy(), c1)
    assert np.array_equal(ist2.cpu().numpy(), ist1)
    ctx.close()

    R = ref.RefKPP(m) if ref.available(name) else None
    if R is None:
        pytest.skip("oracle/_ref not built")
    c_r, ist_r, _, e_r = R.integrate(cf, rf, hf, 0.0, rd[0][0].dt, icf, rcf)
    nv = m.nvar
    sc = np.abs(c1[:nv] - c_r[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(c_r[:nv]))
    print("  replay from the files: GPU bitwise the original step; vs reference max scaled "
          "error %.3g" % sc.max(), flush=True)
    assert (e_r == 1).all()
    assert sc.max() <= 1.0
    assert np.abs(ist_r[2].astype(int) - ist1[2]).max() <= 2
