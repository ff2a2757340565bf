"""GPU tests of the persistent solver's configurations, its fused step (retry,
clip, next Hstart in one launch), the autotuner, and the mixed-mechanism
batch with HIP-graph replay, all through the C ABI.

Oracles: the C oracle (Rodas3 restatement pinned to the reference) for the
integration; the plain GPU step, itself checked against the reference's
golden step outputs, for the step semantics.  The per-cell computation does
not depend on which lane or wave solves a cell, so the persistent variants,
the batch and its graph replays must give bitwise the same results as each
other; against the oracle the bar is the repo-tolerance scaled error <= 1
(ATOL 1e-2, RTOL 0.5e-2) and step counts within 2.
"""
import numpy as np
import pytest

from gckpp_amd import options

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def carbon():
    from gckpp_amd import mechanism
    return mechanism.load("carbon")


def _t(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype or torch.float64)


def _scaled(m, c, ref):
    nv = m.nvar
    return np.abs(c[:nv] - ref[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(ref[:nv]))


@pytest.fixture(scope="module")
def box2k(fullchem):
    """2,000 box cells (1,000 dark, 1,000 lit) with the reference's rates and
    the C oracle's Integrate(0,1200) result."""
    from gckpp_amd import synth
    from oracle import orc, ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    w = synth.box_cells(fullchem, 4096, 0).subset(np.r_[0:1000, 2048:3048])
    rct = ref.RefKPP(fullchem).rates(w)
    ic, rc = options.fullchem_icntrl()
    c_o, i_o, r_o, e_o = orc.Oracle(fullchem).integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc,
                                                         nthreads=16)
    return dict(w=w, rct=rct, ic=ic, rc=rc, c=c_o, ist=i_o, ierr=e_o)


@pytest.mark.parametrize("variant,wpb", [(0, 1), (1, 2), (2, 1), (3, 1), (4, 4), (5, 2), (6, 1), (7, 1),
                                         (8, 4), (9, 2)])
def test_persistent_variants_match_oracle(fullchem, torch_dev, box2k, variant, wpb):
    print("case: persistent variant %d, %d waves per workgroup, 2000 box cells" % (variant, wpb),
          flush=True)
    from gckpp_amd import hip
    b = box2k
    ctx = hip.Context(fullchem, 4096)
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_persistent_config(variant, wpb)
    cfg = ctx.persistent_config()
    assert cfg["variant"] == variant and cfg["waves_per_group"] == wpb
    conc = _t(torch_dev, b["w"].conc)
    ist, rst, ierr = ctx.integrate(conc, _t(torch_dev, b["rct"]), _t(torch_dev, b["w"].hstart),
                                   0.0, 1200.0, b["ic"], b["rc"])
    assert np.array_equal(ierr.cpu().numpy(), b["ierr"])
    assert _scaled(fullchem, conc.cpu().numpy(), b["c"]).max() <= 1.0
    assert np.abs(ist.cpu().numpy()[2] - b["ist"][2]).max() <= 2
    ctx.close()


def test_persistent_step_equals_plain_step_with_retry(fullchem, torch_dev, box2k):
    """The fused persistent step against the plain step (rates, Integrate,
    retry with Hstart=0 from the original C, clip, next Hstart).  A tiny
    per-cell Hstart (1e-13 s) and a step limit (ICNTRL(4)) near the median
    cold-start step count make the first Integrate fail (IERR -6) for many
    cells: of those, the ones whose cold start fits the limit succeed on the
    retry (and then match the oracle's cold-start result), the others fail
    twice -- the same cells in both paths."""
    print("case: persistent step vs plain step, forced retries", flush=True)
    from gckpp_amd import hip
    b = box2k
    w = b["w"]
    ic = b["ic"].copy()
    ic[3] = int(np.median(b["ist"][2])) + 2          # ICNTRL(4) = Max_no_steps
    hs = np.full(w.ncell, 1e-13)
    # first Integrate alone: which cells fail before any retry
    ctx = hip.Context(fullchem, 4096)
    conc = _t(torch_dev, w.conc)
    _, _, e_first = ctx.integrate(conc, _t(torch_dev, b["rct"]), _t(torch_dev, hs), 0.0, 1200.0,
                                  ic, b["rc"])
    e_first = e_first.cpu().numpy()
    ctx.close()
    out = {}
    for solver in (hip.SOLVER_PLAIN, hip.SOLVER_PERSISTENT):
        ctx = hip.Context(fullchem, 4096)
        ctx.set_solver(solver)
        conc = _t(torch_dev, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(torch_dev, w.met), _t(torch_dev, w.photol),
                                                _t(torch_dev, w.het), conc, _t(torch_dev, hs),
                                                1200.0, ic, b["rc"])
        out[solver] = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(),
                       ierr.cpu().numpy(), ok)
        ctx.close()
    cp, hp, ip, ep, okp = out[hip.SOLVER_PLAIN]
    cq, hq, iq, eq, okq = out[hip.SOLVER_PERSISTENT]
    retried_ok = (e_first == -6) & (eq == 1)
    twice = eq < 0
    print("first-attempt failures", int((e_first < 0).sum()), "retried ok", int(retried_ok.sum()),
          "failed twice", int(twice.sum()))
    assert retried_ok.sum() > 50 and twice.sum() > 50
    assert not okp and not okq                      # some cells failed twice
    assert np.array_equal(ep, eq)
    assert set(np.unique(eq[twice])) == {-6}
    good = eq == 1
    assert _scaled(fullchem, cq[:, good], cp[:, good]).max() <= 1.0
    assert np.abs(ip[2][good] - iq[2][good]).max() <= 2
    assert np.all(np.abs(hq[good] - hp[good]) <= 0.5 * hp[good])
    # a retried cell ran the cold start: the oracle's cold-start steps and result
    assert np.abs(iq[2][retried_ok] - b["ist"][2][retried_ok]).max() <= 2
    assert _scaled(fullchem, cq[:, retried_ok], np.maximum(b["c"][:, retried_ok], 0.0)).max() <= 1.0
    assert (cq >= 0).all()


def test_autotuner_picks_a_configuration(fullchem, torch_dev, box2k):
    print("case: autotuner on 2000 box cells", flush=True)
    from gckpp_amd import hip
    b = box2k
    ctx = hip.Context(fullchem, 4096)
    conc = _t(torch_dev, b["w"].conc)
    keep = conc.clone()
    ms = ctx.autotune(conc, _t(torch_dev, b["rct"]), _t(torch_dev, b["w"].hstart), 0.0, 1200.0,
                      b["ic"], b["rc"])
    print("autotune ms [variant][log2 wpb]:\n", ms, ctx.persistent_config())
    assert torch_dev.equal(conc, keep)
    ok = ms[ms > 0]
    assert ok.size >= 5
    cfg = ctx.persistent_config()
    assert abs(ms[cfg["variant"], int(np.log2(cfg["waves_per_group"]))] - ok.min()) < 1e-6
    ctx.close()


def _mixed_inputs(fullchem, hg, carbon, n=(600, 250, 150)):
    from gckpp_amd import synth
    from gckpp_amd.options import icntrl_for
    parts = []
    for m, k, seed in ((fullchem, n[0], 0), (hg, n[1], 1), (carbon, n[2], 2)):
        w = synth.box_cells(m, 2 * k, seed)
        ic, rc = icntrl_for(m.name)
        parts.append((m, w, ic, rc))
    return parts


def test_mixed_batch_equals_per_mechanism_steps_and_graph_replays(fullchem, hg, carbon, torch_dev):
    """One persistent launch over fullchem, Hg and carbon cells against each
    mechanism's plain step; then a captured graph replayed three times (with
    eager work and synchronisation between replays) against three eager
    batch steps: bitwise equal."""
    print("case: mixed batch (600 fullchem, 250 Hg, 150 carbon) and graph replay", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    inputs = _mixed_inputs(fullchem, hg, carbon)
    ctxs = [hip.Context(m, 2048) for m, _, _, _ in inputs]
    # reference: per-mechanism plain steps
    ref = []
    for ctx, (m, w, ic, rc) in zip(ctxs, inputs):
        conc = _t(T, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, w.hstart), 1200.0, ic, rc)
        assert ok
        ref.append((conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy()))
    B = hip.Batch(*ctxs)

    def make_state():
        st = []
        for m, w, ic, rc in inputs:
            d = dict(met=_t(T, w.met), photol=_t(T, w.photol), het=_t(T, w.het), conc=_t(T, w.conc),
                     hstart=_t(T, w.hstart),
                     istatus=T.zeros((20, w.ncell), dtype=T.int32, device="cuda"),
                     ierr=T.zeros((w.ncell,), dtype=T.int32, device="cuda"))
            st.append(d)
        parts = [B.part(icntrl=inp[2], rcntrl=inp[3], **d) for inp, d in zip(inputs, st)]
        return st, parts

    st, parts = make_state()
    B.step(parts, 1200.0)
    T.cuda.synchronize()
    for (m, w, ic, rc), d, r in zip(inputs, st, ref):
        c = d["conc"].cpu().numpy()
        assert np.array_equal(d["ierr"].cpu().numpy(), r[3])
        if m.name == "carbon":
            assert np.array_equal(c, r[0])
        else:
            assert _scaled(m, c, r[0]).max() <= 1.0
            assert np.abs(d["istatus"].cpu().numpy()[2] - r[2][2]).max() <= 2
            assert np.all(np.abs(d["hstart"].cpu().numpy() - r[1]) <= 0.5 * r[1])
    # eager: three consecutive steps from the initial state
    st_e, parts_e = make_state()
    for _ in range(3):
        B.step(parts_e, 1200.0)
    T.cuda.synchronize()
    eager = [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy()) for d in st_e]
    # graph: capture one step on fresh state, replay three times
    st_g, parts_g = make_state()
    B.capture(parts_g, 1200.0)
    for k in range(3):
        B.replay()
        T.cuda.synchronize()
        _ = (st_g[0]["conc"] * 1.0).sum().item()      # eager work between replays
    for (ce, he), d in zip(eager, st_g):
        assert np.array_equal(d["conc"].cpu().numpy(), ce)
        assert np.array_equal(d["hstart"].cpu().numpy(), he)
    B.close()
    for c in ctxs:
        c.close()


def test_mixed_batch_wave_cap_bitwise(fullchem, hg, carbon, torch_dev):
    """The batch's launch with a cap of 8 waves (work stealing over many
    cells per lane) gives the same bits as every resident wave, eager and
    replayed; a cap smaller than a workgroup is refused."""
    print("case: mixed batch with a wave cap", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    inputs = _mixed_inputs(fullchem, hg, carbon)
    ctxs = [hip.Context(m, 2048) for m, _, _, _ in inputs]
    B = hip.Batch(*ctxs)

    def run(cap, graph):
        B.set_waves(cap)
        st = []
        for m, w, ic, rc in inputs:
            st.append(dict(met=_t(T, w.met), photol=_t(T, w.photol), het=_t(T, w.het),
                           conc=_t(T, w.conc), hstart=_t(T, w.hstart),
                           istatus=T.zeros((20, w.ncell), dtype=T.int32, device="cuda"),
                           ierr=T.zeros((w.ncell,), dtype=T.int32, device="cuda")))
        parts = [B.part(icntrl=inp[2], rcntrl=inp[3], **d) for inp, d in zip(inputs, st)]
        if graph:
            B.capture(parts, 1200.0)
            B.replay()
        else:
            B.step(parts, 1200.0)
        T.cuda.synchronize()
        return [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy(), d["istatus"].cpu().numpy(),
                 d["ierr"].cpu().numpy()) for d in st]

    full = run(0, False)
    for cap, graph in ((8, False), (8, True)):
        got = run(cap, graph)
        for a, b in zip(full, got):
            for x, y in zip(a, b):
                assert np.array_equal(x, y)
    B.set_config(4, 4)
    with pytest.raises(hip.GckppError, match="waves"):
        B.set_waves(2)
    B.close()
    for c in ctxs:
        c.close()


@pytest.mark.parametrize("variant,wpb", [(7, 1), (8, 1), (8, 4), (9, 2)])
@pytest.mark.parametrize("name,forced", [("fullchem", False), ("fullchem", True), ("Hg", False),
                                         ("carbon", False)])
def test_step_bitwise_equals_lane_per_cell(name, forced, variant, wpb, mech_of, torch_dev):
    """The cell-per-wave solver (state in LDS, sums shared by the wave's
    lanes) and the lane solver with the LDS operand cache perform the plain
    lane-per-cell solver's operations in the same order: the fused step
    (rates, Integrate, retry, clip, next Hstart) gives bitwise the same
    concentrations, ISTATUS, IERR and Hstart.  forced: Hstart 1e-13 s and
    ICNTRL(4) = 40 make many cells fail and take the retry (the cached
    solver then restarts its schedules mid-cell)."""
    print("case: variant %d (%d waves per group) vs lane-per-cell step, %s, forced retries %s"
          % (variant, wpb, name, forced), flush=True)
    from gckpp_amd import hip, synth
    from gckpp_amd.options import icntrl_for
    T = torch_dev
    m = mech_of(name)
    w = synth.box_cells(m, 1024, 5)
    ic, rc = icntrl_for(name)
    hs = w.hstart
    if forced:
        ic = ic.copy()
        ic[3] = 40
        hs = np.full(w.ncell, 1e-13)
    out = {}
    for v, g in ((4, 1), (variant, wpb)):
        ctx = hip.Context(m, 2048)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        ctx.set_persistent_config(v, g)
        cfg = ctx.persistent_config()
        assert cfg["wave_per_cell"] == (v == hip.WAVE_VARIANT)
        assert cfg["cache_slots"] == {8: 16, 9: 32}.get(v, 0)
        conc = _t(T, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, hs), 1200.0, ic, rc)
        out[v] = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy())
        ctx.close()
    a, b = out[4], out[variant]
    if forced:
        print("failed cells", int((a[3] < 0).sum()), "of", w.ncell)
        assert (a[3] < 0).sum() > 20 and (a[3] == 1).sum() > 20
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("variant", [7, 8, 9])
def test_mixed_batch_bitwise_and_graph_replays(variant, fullchem, hg, carbon, torch_dev):
    """The mixed-mechanism batch with the cell-per-wave kernel (7) and the
    cached lane kernels (8, 9): one step bitwise equal to the plain
    lane-per-cell batch step and to each mechanism's own persistent step; a
    captured graph replayed three times equals three eager steps."""
    print("case: mixed batch with persistent variant %d, graph replay" % variant, flush=True)
    from gckpp_amd import hip
    T = torch_dev
    inputs = _mixed_inputs(fullchem, hg, carbon)
    ctxs = [hip.Context(m, 2048) for m, _, _, _ in inputs]
    B = hip.Batch(*ctxs)

    def make_state():
        st = []
        for m, w, ic, rc in inputs:
            st.append(dict(met=_t(T, w.met), photol=_t(T, w.photol), het=_t(T, w.het),
                           conc=_t(T, w.conc), hstart=_t(T, w.hstart),
                           istatus=T.zeros((20, w.ncell), dtype=T.int32, device="cuda"),
                           ierr=T.zeros((w.ncell,), dtype=T.int32, device="cuda")))
        return st, [B.part(icntrl=inp[2], rcntrl=inp[3], **d) for inp, d in zip(inputs, st)]

    res = {}
    for v in (4, variant):
        B.set_config(v, 1)
        st, parts = make_state()
        B.step(parts, 1200.0)
        T.cuda.synchronize()
        res[v] = [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy(),
                   d["istatus"].cpu().numpy(), d["ierr"].cpu().numpy()) for d in st]
    for ra, rb in zip(res[4], res[variant]):
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
    # each mechanism alone through its own context's persistent step
    for ctx, (m, w, ic, rc), rb in zip(ctxs, inputs, res[variant]):
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        ctx.set_persistent_config(4, 1)
        conc = _t(T, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, w.hstart), 1200.0, ic, rc)
        assert np.array_equal(conc.cpu().numpy(), rb[0]), m.name
        assert np.array_equal(hnew.cpu().numpy(), rb[1]), m.name
    B.set_config(variant, 1)
    st_e, parts_e = make_state()
    for _ in range(3):
        B.step(parts_e, 1200.0)
    T.cuda.synchronize()
    eager = [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy()) for d in st_e]
    st_g, parts_g = make_state()
    B.capture(parts_g, 1200.0)
    for k in range(3):
        B.replay()
        T.cuda.synchronize()
        _ = (st_g[0]["conc"] * 1.0).sum().item()
    for (ce, he), d in zip(eager, st_g):
        assert np.array_equal(d["conc"].cpu().numpy(), ce)
        assert np.array_equal(d["hstart"].cpu().numpy(), he)
    B.close()
    for c in ctxs:
        c.close()


def test_order_by_h0_is_a_stable_bin_sort(torch_dev):
    """gckpp_order_by_h0_dev: a permutation of the cells, bins of log10(h0)
    (quarter decades from 1e-5 s) nondecreasing along it, increasing cell
    index within a bin; h0 <= 0 last."""
    print("case: order by predicted h0, 50,000 cells", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    rng = np.random.default_rng(5)
    n = 50000
    width = (5.0 + np.log10(1200.0)) / 32
    b = rng.integers(0, 32, n)
    h0 = 10.0 ** (-5.0 + (b + rng.uniform(0.2, 0.8, n)) * width)   # away from the bin edges
    h0[rng.integers(0, n, 300)] = 0.0
    order = hip.order_by_h0(_t(T, h0)).cpu().numpy()
    assert np.array_equal(np.sort(order), np.arange(n))
    key = np.where(h0 > 0, np.clip(np.floor((np.log10(np.where(h0 > 0, h0, 1.0)) + 5.0) / width), 0, 31), 31)
    k = key[order]
    assert (np.diff(k) >= 0).all()
    same = np.diff(k) == 0
    assert (np.diff(order)[same] > 0).all()


@pytest.mark.parametrize("name,forced", [("fullchem", False), ("fullchem", True), ("Hg", False)])
def test_split_persistent_step_bitwise(name, forced, mech_of, torch_dev):
    """The split persistent step -- the first cells of the h0 order on the
    cell-per-wave kernel on a second stream, the rest on the lane kernel at
    the same time -- gives bitwise the single-kernel step's results, also
    when cells fail and take the retry."""
    print("case: split persistent step, %s, forced retries %s" % (name, forced), flush=True)
    from gckpp_amd import hip, synth
    from gckpp_amd.options import icntrl_for
    T = torch_dev
    m = mech_of(name)
    w = synth.box_cells(m, 1024, 9)
    ic, rc = icntrl_for(name)
    hs = w.hstart
    if forced:
        ic = ic.copy()
        ic[3] = 40
        hs = np.full(w.ncell, 1e-13)
    rng = np.random.default_rng(2)
    h0_guess = _t(T, 10.0 ** rng.uniform(-5, 3, w.ncell))
    order = hip.order_by_h0(h0_guess)
    out = {}
    for split in (0, 150):
        ctx = hip.Context(m, 2048)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        ctx.set_cell_order(order)
        ctx.set_persistent_split(split)
        conc = _t(T, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, hs), 1200.0, ic, rc)
        out[split] = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy())
        # Integrate alone, split the same way
        c2 = _t(T, w.conc)
        rct = ctx.rates(_t(T, w.met), _t(T, w.photol), _t(T, w.het), c2)
        ist2, rst2, ierr2 = ctx.integrate(c2, rct, _t(T, hs), 0.0, 1200.0, ic, rc)
        out[split] += (c2.cpu().numpy(), ist2.cpu().numpy(), ierr2.cpu().numpy())
        ctx.close()
    if forced:
        assert (out[0][3] < 0).sum() > 20 and (out[0][3] == 1).sum() > 20
    for x, y in zip(out[0], out[150]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("name,forced,variant", [("fullchem", False, 4), ("fullchem", True, 4),
                                                 ("fullchem", False, 8), ("Hg", False, 1)])
def test_handoff_persistent_step_bitwise(name, forced, variant, mech_of, torch_dev):
    """The tail hand-off -- the lane kernel's last unfinished cells saved at
    a step boundary and continued by the cell-per-wave kernel -- gives
    bitwise the results without it, for the step (with forced retries, so
    that cells are handed off in their second integration too) and for
    Integrate alone, also together with the split solve; the hand-off list
    is used and never overflows; the step-size history log is the same too
    (the cell-per-wave kernel continues a handed-off cell's rows)."""
    print("case: hand-off, %s, forced retries %s, variant %d" % (name, forced, variant), flush=True)
    from gckpp_amd import hip, synth
    from gckpp_amd.options import icntrl_for
    T = torch_dev
    m = mech_of(name)
    w = synth.box_cells(m, 1024, 9)
    ic, rc = icntrl_for(name)
    hs = w.hstart
    if forced:
        ic = ic.copy()
        ic[3] = 40
        hs = np.full(w.ncell, 1e-13)
    order = hip.order_by_h0(_t(T, 10.0 ** np.random.default_rng(2).uniform(-5, 3, w.ncell)))
    out, counts = {}, {}
    for split, hand in ((0, 0), (0, 400), (150, 400), (0, 1024)):
        ctx = hip.Context(m, 2048)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        ctx.set_persistent_config(variant, 1)
        ctx.set_cell_order(order)
        ctx.set_persistent_split(split)
        ctx.set_persistent_handoff(hand)
        hlog = T.zeros((48, w.ncell), dtype=T.float32, device="cuda")
        ctx.set_step_log(hlog, 24)
        conc = _t(T, w.conc)
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, hs), 1200.0, ic, rc)
        ctx.set_step_log(None)
        res = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy(),
               hlog.cpu().numpy())
        stats = ctx.handoff_stats()
        counts[(split, hand)] = stats[0]
        if hand:   # the cell-per-wave launch attempted the handed-off cells; both launches timed
            assert stats[0] > 0 and stats[1] > 0 and stats[2] > 0 and stats[3] > 0, stats
        c2 = _t(T, w.conc)
        rct = ctx.rates(_t(T, w.met), _t(T, w.photol), _t(T, w.het), c2)
        ist2, rst2, ierr2 = ctx.integrate(c2, rct, _t(T, hs), 0.0, 1200.0, ic, rc)
        out[(split, hand)] = res + (c2.cpu().numpy(), ist2.cpu().numpy(), rst2.cpu().numpy(),
                                    ierr2.cpu().numpy())
        ctx.close()
    print("handed off:", counts, flush=True)
    if forced:
        assert (out[(0, 0)][3] < 0).sum() > 20 and (out[(0, 0)][3] == 1).sum() > 20
    assert counts[(0, 0)] == 0
    for key in ((0, 400), (150, 400), (0, 1024)):
        assert 0 < counts[key] <= key[1]
        for x, y in zip(out[(0, 0)], out[key]):
            assert np.array_equal(x, y), key
    # every cell unfinished from the start: all handed off at their first step
    assert counts[(0, 1024)] == 1024


@pytest.mark.parametrize("hand", [700, 2000])
def test_mixed_batch_handoff_bitwise_and_graph_replays(hand, fullchem, hg, carbon, torch_dev):
    """The tail hand-off in the mixed-mechanism batch (fullchem and Hg cells
    continued by the mixed cell-per-wave kernel, carbon cells finishing on
    their lanes): one step bitwise equal to the batch step without it; a
    captured graph with the hand-off replayed three times equals three eager
    steps without it."""
    print("case: mixed batch, hand-off %d of 2000 cells, graph replay" % hand, flush=True)
    from gckpp_amd import hip
    T = torch_dev
    inputs = _mixed_inputs(fullchem, hg, carbon)
    ctxs = [hip.Context(m, 2048) for m, _, _, _ in inputs]
    B = hip.Batch(*ctxs)
    B.set_config(4, 1)

    def make_state():
        st = []
        for m, w, ic, rc in inputs:
            st.append(dict(met=_t(T, w.met), photol=_t(T, w.photol), het=_t(T, w.het),
                           conc=_t(T, w.conc), hstart=_t(T, w.hstart),
                           istatus=T.zeros((20, w.ncell), dtype=T.int32, device="cuda"),
                           ierr=T.zeros((w.ncell,), dtype=T.int32, device="cuda")))
        return st, [B.part(icntrl=inp[2], rcntrl=inp[3], **d) for inp, d in zip(inputs, st)]

    def fetch(st):
        return [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy(), d["istatus"].cpu().numpy(),
                 d["ierr"].cpu().numpy()) for d in st]

    res = {}
    for h in (0, hand):
        B.set_handoff(h)
        st, parts = make_state()
        B.step(parts, 1200.0)
        T.cuda.synchronize()
        res[h] = fetch(st)
        if h:
            n = B.handoff_count()
            print("handed off", n, flush=True)
            assert 0 < n <= h
    for ra, rb in zip(res[0], res[hand]):
        for x, y in zip(ra, rb):
            assert np.array_equal(x, y)
    B.set_handoff(0)
    st_e, parts_e = make_state()
    for _ in range(3):
        B.step(parts_e, 1200.0)
    T.cuda.synchronize()
    eager = [(d["conc"].cpu().numpy(), d["hstart"].cpu().numpy()) for d in st_e]
    B.set_handoff(hand)
    st_g, parts_g = make_state()
    B.capture(parts_g, 1200.0)
    for k in range(3):
        B.replay()
        T.cuda.synchronize()
        _ = (st_g[0]["conc"] * 1.0).sum().item()
    for (ce, he), d in zip(eager, st_g):
        assert np.array_equal(d["conc"].cpu().numpy(), ce)
        assert np.array_equal(d["hstart"].cpu().numpy(), he)
    # a larger hand-off reallocates the list the captured graph points at:
    # the graph is dropped and replay refuses (before any launch) until the
    # next capture
    B.set_handoff(4 * hand)
    with pytest.raises(hip.GckppError, match="nothing captured"):
        B.replay()
    B.capture(parts_g, 1200.0)
    B.replay()
    T.cuda.synchronize()
    assert np.isfinite(st_g[0]["conc"].cpu().numpy()).all()
    B.close()
    for c in ctxs:
        c.close()
