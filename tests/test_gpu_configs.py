"""The five benchmark workloads at their full sizes, through the C ABI.

  C1  4,096 synthetic box cells (seed 0), plain and persistent steps
  C2  the 72x46x72 grid (238,464 cells), plain and persistent steps
  C3  the 144x91x72 grid (943,488 cells) of fullchem, Hg and carbon cells in
      one persistent launch per step: 24 steps replayed from one captured
      HIP graph with a synthetic emission/transport perturbation between
      steps, against the same 24 steps run eagerly
  C4  the committed step-size network on the C2 grid (learned h0, cells in
      cost order, persistent solver)
  C5  the C90-sized workload, six 90x90x72 faces (3,499,200 cells) through
      one context sized for one face, learned h0 + persistent solver

On every cell: IERR = 1, finite concentrations, C >= 0 after the step's
clip (fullchem, Hg; the carbon caller keeps negatives), the next Hstart in
(0, dt].  On a strided sample of each workload (about 4,000 cells per
mechanism): the compiled reference (oracle/_ref, the reference's own KPP
code driven as Do_FullChem drives it: rates, Integrate, retry, clip) on the
same inputs and the same per-cell Hstart -- for the learned legs that is
the predicted h0, an RCNTRL(3) the reference accepts like any other --
must agree to the repo tolerance: |c - c_ref| <= ATOL + RTOL |c_ref| for
every variable species (scaled error <= 1) and step counts within 2.
Reference loops: GeosCore/fullchem_mod.F90:528-1551,
GeosCore/mercury_mod.F90:1132-1167, GeosCore/carbon_gases_mod.F90:450-604.
"""
import numpy as np
import pytest

from gckpp_amd import options

pytestmark = pytest.mark.gpu

DT = 1200.0


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def refkpp():
    from oracle import ref
    cache = {}

    def get(m):
        if not ref.available(m.name):
            pytest.skip("oracle/_ref not built")
        if m.name not in cache:
            cache[m.name] = ref.RefKPP(m)
        return cache[m.name]
    return get


def _dev(T, a, dtype=None):
    return T.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype or T.float64)


def _check_every_cell(T, m, conc, hnew, ierr, label, clip=True):
    """Properties of a step's result on every cell, evaluated on the device."""
    bad = int((ierr != 1).sum().item())
    assert bad == 0, "%s: %d cells with IERR != 1" % (label, bad)
    assert bool(T.isfinite(conc).all().item()), "%s: non-finite concentrations" % label
    if clip:
        assert bool((conc >= 0).all().item()), "%s: negative concentrations after the clip" % label
    if hnew is not None:
        ok = (hnew > 0) & (hnew <= DT)
        assert bool(ok.all().item()), "%s: next Hstart outside (0, dt]" % label


def _vs_reference(R, m, w_sample, c_gpu, st_gpu, label, hstart=None):
    """The compiled reference's step on the sampled cells (hstart: per-cell
    RCNTRL(3), default the workload's) against the GPU result there."""
    ic, rc = options.icntrl_for(m.name)
    ws = w_sample
    if hstart is not None:
        ws.hstart = np.ascontiguousarray(hstart, dtype=np.float64)
    c_r, st_r, h_r, e_r = R.fullchem_step(ws, ic, rc, nthreads=16)
    nv = m.nvar
    sc = np.abs(c_gpu[:nv] - c_r[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(c_r[:nv]))
    dsteps = np.abs(st_gpu[2].astype(np.int64) - st_r[2]) if st_gpu is not None else np.zeros(1)
    print("  %s vs reference on %d cells: max scaled error %.3g (median %.3g), max step-count "
          "difference %d" % (label, ws.ncell, sc.max(), np.median(sc), dsteps.max()), flush=True)
    assert (e_r == 1).all()
    assert sc.max() <= 1.0, label
    assert dsteps.max() <= 2, label
    return sc


def _strided(n, k):
    return np.unique(np.linspace(0, n - 1, k).astype(np.int64))


def test_c1_golden_step_outputs(fullchem, T):
    """The fullchem golden step (the compiled reference's Do_FullChem step on
    eight box cells: rates, Integrate, retry, clip, KPPHvalue) against the
    GPU step, plain and persistent: same IERR, concentrations to 1e-3 of the
    tolerance, the next Hstart (RSTATUS(3)) to 1e-6 relative."""
    from pathlib import Path
    from gckpp_amd import hip
    g = dict(np.load(Path(__file__).resolve().parent / "golden" / "fullchem_box.npz"))
    ic, rc = options.fullchem_icntrl()
    n = g["conc"].shape[1]
    for solver, name in ((hip.SOLVER_PLAIN, "plain"), (hip.SOLVER_PERSISTENT, "persistent")):
        print("case: golden step, %s solver" % name, flush=True)
        ctx = hip.Context(fullchem, 64)
        ctx.set_solver(solver)
        conc = _dev(T, g["conc"])
        hnew, ist, ierr, ok = ctx.fullchem_step(_dev(T, g["met"]), _dev(T, g["photol"]), _dev(T, g["het"]),
                                                conc, T.zeros(n, dtype=T.float64, device="cuda"), DT,
                                                ic, rc)
        assert ok
        c = conc.cpu().numpy()
        nv = fullchem.nvar
        sc = np.abs(c[:nv] - g["step_conc"][:nv]) / (1e-2 + 5e-3 * np.abs(g["step_conc"][:nv]))
        dh = np.abs(hnew.cpu().numpy() - g["step_hnew"]) / g["step_hnew"]
        print("  max scaled error %.3g, max relative Hnew difference %.3g" % (sc.max(), dh.max()))
        assert np.array_equal(ierr.cpu().numpy(), g["step_ierr"])
        assert sc.max() < 1e-3
        assert dh.max() < 1e-6
        assert np.array_equal(c[nv:], g["step_conc"][nv:])      # fixed species pass through
        ctx.close()


@pytest.mark.parametrize("solver", ["plain", "persistent"])
def test_c1_all_box_cells(solver, fullchem, T, refkpp):
    """C1 on all 4,096 box cells: every cell's step against the compiled
    reference's step on the same inputs."""
    print("case: C1, 4096 box cells, %s step" % solver, flush=True)
    from gckpp_amd import hip, synth
    w = synth.box_cells(fullchem, 4096, 0)
    ic, rc = options.fullchem_icntrl()
    ctx = hip.Context(fullchem, w.ncell)
    ctx.set_solver(hip.SOLVER_PLAIN if solver == "plain" else hip.SOLVER_PERSISTENT)
    if solver == "persistent":
        ctx.set_persistent_handoff(512)
    conc = _dev(T, w.conc)
    hnew, ist, ierr, ok = ctx.fullchem_step(_dev(T, w.met), _dev(T, w.photol), _dev(T, w.het), conc,
                                            _dev(T, w.hstart), DT, ic, rc)
    T.cuda.synchronize()
    _check_every_cell(T, fullchem, conc, hnew, ierr, "C1 " + solver)
    _vs_reference(refkpp(fullchem), fullchem, w, conc.cpu().numpy(), ist.cpu().numpy(), "C1 " + solver)
    ctx.close()


@pytest.fixture(scope="module")
def grid72(fullchem, T):
    """The 72x46x72 grid (238,464 cells, seed 1, 12 UTC) on the device, a
    context for all of it, and a strided sample of 4,096 cells."""
    from gckpp_amd import hip, synth
    w = synth.grid_cells(fullchem, 72, 46, 72, seed=1)
    d = dict(met=_dev(T, w.met), photol=_dev(T, w.photol), het=_dev(T, w.het), conc0=_dev(T, w.conc))
    ctx = hip.Context(fullchem, w.ncell)
    idx = _strided(w.ncell, 4096)
    yield dict(w=w, d=d, ctx=ctx, idx=idx)
    ctx.close()
    T.cuda.empty_cache()


@pytest.mark.parametrize("solver", ["plain", "persistent"])
def test_c2_grid_step(solver, grid72, fullchem, T, refkpp):
    """C2: the cold 1200 s step of the 72x46x72 grid (plain one-lane-per-cell
    solver; persistent solver with the tail hand-off of the benchmark)."""
    print("case: C2, 72x46x72 grid (238464 cells), %s step" % solver, flush=True)
    from gckpp_amd import hip
    g = grid72
    w, d, ctx, idx = g["w"], g["d"], g["ctx"], g["idx"]
    ic, rc = options.fullchem_icntrl()
    ctx.set_solver(hip.SOLVER_PLAIN if solver == "plain" else hip.SOLVER_PERSISTENT)
    ctx.set_persistent_handoff(32768 if solver == "persistent" else 0)
    conc = d["conc0"].clone()
    cold = T.zeros(w.ncell, dtype=T.float64, device="cuda")
    hnew, ist, ierr, ok = ctx.fullchem_step(d["met"], d["photol"], d["het"], conc, cold, DT, ic, rc)
    T.cuda.synchronize()
    _check_every_cell(T, fullchem, conc, hnew, ierr, "C2 " + solver)
    print("  mean attempts per cell %.2f" % ist[2].double().mean().item())
    _vs_reference(refkpp(fullchem), fullchem, w.subset(idx), conc[:, idx].cpu().numpy(),
                  ist[:, idx].cpu().numpy(), "C2 " + solver)
    ctx.set_persistent_handoff(0)


def test_c4_learned_h0_on_grid(grid72, fullchem, T, refkpp):
    """C4: the committed network's h0 and cost order on the C2 grid, the
    persistent solver with the tail hand-off.  The reference, started from
    the same h0 (its RCNTRL(3)), gives the same result on the sample: the
    network only chooses the first step, which the controller then judges."""
    print("case: C4, learned h0 on the 72x46x72 grid", flush=True)
    from gckpp_amd import hip, learned
    path = learned.default_model_path("fullchem")
    if not path.exists():
        pytest.skip("no committed model")
    g = grid72
    w, d, ctx, idx = g["w"], g["d"], g["ctx"], g["idx"]
    ic, rc = options.fullchem_icntrl()
    model = learned.StepModel(fullchem).load(path)
    conc = d["conc0"].clone()
    h0, cls = model.predict(d["met"], d["photol"], conc, dt=DT)
    assert bool(((h0 >= 1e-5) & (h0 <= DT)).all().item())
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_persistent_handoff(32768)
    ctx.set_cell_order(hip.order_by_class(cls))
    hnew, ist, ierr, ok = ctx.fullchem_step(d["met"], d["photol"], d["het"], conc, h0.clone(), DT, ic, rc)
    T.cuda.synchronize()
    ctx.set_cell_order(None)
    ctx.set_persistent_handoff(0)
    _check_every_cell(T, fullchem, conc, hnew, ierr, "C4 learned")
    print("  mean attempts per cell %.2f" % ist[2].double().mean().item())
    _vs_reference(refkpp(fullchem), fullchem, w.subset(idx), conc[:, idx].cpu().numpy(),
                  ist[:, idx].cpu().numpy(), "C4 learned", hstart=h0[idx].cpu().numpy())


def test_c3_mixed_grid_24_graph_replays(fullchem, hg, T, refkpp):
    """C3: 943,488 cells, fullchem/Hg/carbon 70/20/10 at random, one
    persistent launch (with the tail hand-off) per step.  The first step's
    sample against each mechanism's compiled reference; then one captured
    step replayed 24 times with a perturbation between replays, bitwise
    equal after every step to the same 24 steps run eagerly on a copy."""
    print("case: C3, 144x91x72 mixed grid, 24 graph replays", flush=True)
    from gckpp_amd import hip, mechanism, synth
    mechs = [fullchem, hg, mechanism.load("carbon")]
    parts, _ = synth.mixed_grid(mechs, 144, 91, 72, seed=3)
    print("  cells per mechanism", [p.ncell for p in parts], flush=True)
    ctxs = [hip.Context(m, p.ncell) for m, p in zip(mechs, parts)]
    B = hip.Batch(*ctxs)
    B.set_handoff(32768)

    def state():
        st = []
        for m, p in zip(mechs, parts):
            ic, rc = options.icntrl_for(m.name)
            st.append(dict(met=_dev(T, p.met), photol=_dev(T, p.photol), het=_dev(T, p.het),
                           conc=_dev(T, p.conc), hstart=T.zeros(p.ncell, dtype=T.float64, device="cuda"),
                           istatus=T.zeros((20, p.ncell), dtype=T.int32, device="cuda"),
                           ierr=T.zeros(p.ncell, dtype=T.int32, device="cuda"), ic=ic, rc=rc))
        return st

    def parts_of(st):
        return [B.part(met=s["met"], photol=s["photol"], het=s["het"], conc=s["conc"], hstart=s["hstart"],
                       icntrl=s["ic"], rcntrl=s["rc"], istatus=s["istatus"], ierr=s["ierr"]) for s in st]

    eager, graph = state(), state()
    pe, pg = parts_of(eager), parts_of(graph)
    B.capture(pg, DT)
    ge = T.Generator(device="cuda").manual_seed(11)
    gg = T.Generator(device="cuda").manual_seed(11)
    for k in range(24):
        if k > 0:
            for m, se, sg in zip(mechs, eager, graph):
                synth.perturb_dev(m, se["conc"], se["met"], ge)
                synth.perturb_dev(m, sg["conc"], sg["met"], gg)
        B.step(pe, DT)
        B.replay()
        T.cuda.synchronize()
        for m, se, sg in zip(mechs, eager, graph):
            lab = "C3 step %d %s" % (k + 1, m.name)
            _check_every_cell(T, m, sg["conc"], sg["hstart"] if m.name != "carbon" else None,
                              sg["ierr"], lab, clip=m.name != "carbon")
            assert T.equal(se["conc"], sg["conc"]), lab + ": replay differs from the eager step"
            assert T.equal(se["hstart"], sg["hstart"]), lab
            assert T.equal(se["istatus"], sg["istatus"]), lab
        if k == 0:
            for m, p, s in zip(mechs, parts, graph):
                idx = _strided(p.ncell, 4096)
                _vs_reference(refkpp(m), m, p.subset(idx), s["conc"][:, idx].cpu().numpy(),
                              s["istatus"][:, idx].cpu().numpy() if m.name != "carbon" else None,
                              "C3 first step, %s" % m.name)
        if k % 6 == 5:
            print("  %d replays equal to the eager steps; attempts/cell %s; handed off %d"
                  % (k + 1, [round(float(s["istatus"][2].double().mean().item()), 2) for s in graph],
                     B.handoff_count()), flush=True)
    B.close()
    for c in ctxs:
        c.close()
    T.cuda.empty_cache()


def test_c5_c90_faces_learned_persistent(fullchem, T, refkpp):
    """C5: 6x90x90x72 = 3,499,200 cells as six face-sized chunks through one
    context whose workspace is sized for one face; learned h0, cost order,
    persistent solver with the tail hand-off; the reference from the same
    h0 on 683 strided cells of every face."""
    print("case: C5, six 90x90x72 faces (3499200 cells), learned + persistent", flush=True)
    from gckpp_amd import hip, learned, synth
    path = learned.default_model_path("fullchem")
    if not path.exists():
        pytest.skip("no committed model")
    ic, rc = options.fullchem_icntrl()
    nface = 90 * 90 * 72
    ctx = hip.Context(fullchem, nface)
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_persistent_handoff(32768)
    model = learned.StepModel(fullchem).load(path)
    idx = _strided(nface, 683)
    wsamp, csamp, ssamp, hsamp = [], [], [], []
    for f in range(6):
        w = synth.grid_cells(fullchem, 90, 90, 72, seed=200 + f, hour_utc=(12.0 + 4.0 * f) % 24.0)
        met, photol, het, conc = _dev(T, w.met), _dev(T, w.photol), _dev(T, w.het), _dev(T, w.conc)
        h0, cls = model.predict(met, photol, conc, dt=DT)
        ctx.set_cell_order(hip.order_by_class(cls))
        hnew, ist, ierr, ok = ctx.fullchem_step(met, photol, het, conc, h0.clone(), DT, ic, rc)
        T.cuda.synchronize()
        ctx.set_cell_order(None)
        _check_every_cell(T, fullchem, conc, hnew, ierr, "C5 face %d" % f)
        wsamp.append(w.subset(idx))
        csamp.append(conc[:, idx].cpu().numpy())
        ssamp.append(ist[:, idx].cpu().numpy())
        hsamp.append(h0[idx].cpu().numpy())
        print("  face %d: every cell checked, mean attempts %.2f" % (f, ist[2].double().mean().item()),
              flush=True)
        del met, photol, het, conc, h0, cls, hnew, ist, ierr, w
    ctx.close()
    T.cuda.empty_cache()
    ws = wsamp[0]
    ws = synth.Workload(fullchem, np.concatenate([x.met for x in wsamp], 1),
                        np.concatenate([x.photol for x in wsamp], 1),
                        np.concatenate([x.het for x in wsamp], 1) if ws.het.shape[1] > 1 else ws.het,
                        np.concatenate([x.conc for x in wsamp], 1), np.concatenate([x.hstart for x in wsamp]),
                        ws.dt, "C90 sample")
    _vs_reference(refkpp(fullchem), fullchem, ws, np.concatenate(csamp, 1), np.concatenate(ssamp, 1),
                  "C5 learned", hstart=np.concatenate(hsamp))
