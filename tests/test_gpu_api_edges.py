"""Edge cases of the C ABI on the GPU: empty batches, batches larger than a
context's max_cells, option errors that the reference's Rosenbrock() reports
per call (here per cell: IERR -1, -2, -5 with concentrations untouched),
options it does not implement (an error return with a message), and null
arguments.  The reference codes follow KPP/fullchem/gckpp_Integrator.F90's
option checks (:354-476) and ros_ErrorMsg (:538-575); the option merge
(ICNTRL_U where nonzero, RCNTRL_U where positive) is INTEGRATE's
(:112-117)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _t(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype or torch.float64)


@pytest.fixture(scope="module")
def box(fullchem, torch_dev):
    from gckpp_amd import hip, synth
    from gckpp_amd.options import fullchem_icntrl
    T = torch_dev
    w = synth.box_cells(fullchem, 600, 11)
    ctx = hip.Context(fullchem, 1024)
    rct = ctx.rates(_t(T, w.met), _t(T, w.photol), _t(T, w.het), _t(T, w.conc))
    T.cuda.synchronize()
    ctx.close()
    ic, rc = fullchem_icntrl()
    return dict(w=w, rct=rct, ic=ic, rc=rc)


def test_empty_batch_is_a_no_op(fullchem, torch_dev, box):
    print("case: ncell = 0 through rates, integrate, step", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    w = box["w"]
    ctx = hip.Context(fullchem, 64)
    e = lambda rows: T.empty((rows, 0), dtype=T.float64, device="cuda")
    out = ctx.rates(e(w.met.shape[0]), e(w.photol.shape[0]), e(w.het.shape[0]), e(fullchem.nspec))
    assert out.shape == (fullchem.nreact, 0)
    for solver in (hip.SOLVER_PLAIN, hip.SOLVER_PERSISTENT):
        ctx.set_solver(solver)
        ist, rst, ierr = ctx.integrate(e(fullchem.nspec), e(fullchem.nreact),
                                       T.empty(0, dtype=T.float64, device="cuda"), 0.0, 1200.0,
                                       box["ic"], box["rc"])
        assert ierr.numel() == 0
        hnew, ist, ierr, ok = ctx.fullchem_step(e(w.met.shape[0]), e(w.photol.shape[0]),
                                                e(w.het.shape[0]), e(fullchem.nspec),
                                                T.empty(0, dtype=T.float64, device="cuda"),
                                                1200.0, box["ic"], box["rc"])
        assert ok and ierr.numel() == 0
    ctx.close()


def test_batch_larger_than_max_cells(fullchem, torch_dev, box):
    """Integrate takes any number of cells: the plain solver runs chunks of
    max_cells, the persistent solver streams cells through its resident
    lanes; both give bitwise the result of a context that holds them all."""
    print("case: 600 cells through contexts of 128 and 1024 max cells", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    w = box["w"]
    res = {}
    for maxc, solver in ((1024, hip.SOLVER_PLAIN), (128, hip.SOLVER_PLAIN),
                         (1024, hip.SOLVER_PERSISTENT), (128, hip.SOLVER_PERSISTENT)):
        ctx = hip.Context(fullchem, maxc)
        ctx.set_solver(solver)
        conc = _t(T, w.conc)
        ist, rst, ierr = ctx.integrate(conc, box["rct"], _t(T, w.hstart), 0.0, 1200.0,
                                       box["ic"], box["rc"])
        res[(maxc, solver)] = (conc.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy())
        ctx.close()
    ref = res[(1024, hip.SOLVER_PLAIN)]
    assert (ref[2] == 1).all()
    for k, v in res.items():
        for x, y in zip(ref, v):
            assert np.array_equal(x, y), k
    # the component entry points use the per-cell workspace: more cells than
    # max_cells is an error, reported with a message
    ctx = hip.Context(fullchem, 128)
    with pytest.raises(hip.GckppError, match="exceeds max_cells"):
        ctx.fun(_t(T, w.conc), box["rct"])
    ctx.close()


@pytest.mark.parametrize("case", ["max_steps", "method", "tolerance"])
def test_option_errors_are_per_cell_codes(case, fullchem, torch_dev, box):
    """Rosenbrock() rejects its options before integrating: IERR -1 for a
    negative ICNTRL(4), -2 for an unknown ICNTRL(3), -5 for a relative
    tolerance outside (10*roundoff, 1); every cell reports the code and
    keeps its concentrations."""
    print("case: option error %s" % case, flush=True)
    from gckpp_amd import hip
    T = torch_dev
    w = box["w"]
    ic = box["ic"].copy()
    want = {"max_steps": -1, "method": -2, "tolerance": -5}[case]
    for solver in (hip.SOLVER_PLAIN, hip.SOLVER_PERSISTENT):
        ctx = hip.Context(fullchem, 1024)
        ctx.set_solver(solver)
        if case == "max_steps":
            ic[3] = -4
        elif case == "method":
            ic[2] = 9
        else:
            rtol = fullchem.rtol().copy()
            rtol[5] = 1.0
            ctx.set_tolerances(fullchem.atol(), rtol)
        conc = _t(T, w.conc)
        ist, rst, ierr = ctx.integrate(conc, box["rct"], _t(T, w.hstart), 0.0, 1200.0, ic,
                                       box["rc"])
        assert (ierr.cpu().numpy() == want).all()
        assert (ist.cpu().numpy() == 0).all()
        assert np.array_equal(conc.cpu().numpy(), w.conc)
        ctx.close()


@pytest.mark.parametrize("k,val,msg", [(2, 2, "only Rodas3"), (11, 1, "auto-reduce"),
                                       (14, 5, "frozen")])
def test_unimplemented_options_fail_with_a_message(k, val, msg, fullchem, torch_dev, box):
    """Other Rosenbrock methods (ICNTRL(3) = 1..6 except Rodas3), the
    auto-reduce integrator (ICNTRL(12) = 1) and rate updates inside the step
    (ICNTRL(15) != -1) are outside the path: the call fails and says why."""
    print("case: ICNTRL(%d) = %d" % (k + 1, val), flush=True)
    from gckpp_amd import hip
    T = torch_dev
    w = box["w"]
    ic = box["ic"].copy()
    ic[k] = val
    ctx = hip.Context(fullchem, 1024)
    with pytest.raises(hip.GckppError, match=msg):
        ctx.integrate(_t(T, w.conc), box["rct"], _t(T, w.hstart), 0.0, 1200.0, ic, box["rc"])
    ctx.close()


def test_null_arguments_fail(fullchem, torch_dev, box):
    print("case: null context and null arrays", flush=True)
    from gckpp_amd import hip
    L = hip.load()
    ica = np.ascontiguousarray(box["ic"], dtype=np.int32)
    rca = np.ascontiguousarray(box["rc"], dtype=np.float64)
    ic, rc = ctypes.c_void_p(ica.ctypes.data), ctypes.c_void_p(rca.ctypes.data)
    assert L.gckpp_integrate_dev(None, 10, 0.0, 1200.0, ic, rc, None, None, None, None, None,
                                 None, None) == -1
    assert b"null context" in L.gckpp_last_error()
    ctx = hip.Context(fullchem, 64)
    assert L.gckpp_integrate_dev(ctx.h, 10, 0.0, 1200.0, ic, rc, None, None, None, None, None,
                                 None, None) == -1
    assert b"null conc or rconst" in L.gckpp_last_error()
    assert L.gckpp_integrate_dev(ctx.h, -3, 0.0, 1200.0, ic, rc, None, None, None, None, None,
                                 None, None) == -1
    assert b"negative ncell" in L.gckpp_last_error()
    ctx.close()


def test_step_option_error_same_in_both_solvers(fullchem, torch_dev, box):
    """An option error in the fused step: every cell reports it (after the
    retry fails the same way), ISTATUS is zero, Hnew = RSTATUS(3) = 0, and C
    is only clipped -- the same in the plain and the persistent step."""
    print("case: fullchem step with ICNTRL(4) < 0, plain and persistent", flush=True)
    from gckpp_amd import hip
    T = torch_dev
    w = box["w"]
    ic = box["ic"].copy()
    ic[3] = -4
    out = {}
    for solver in (hip.SOLVER_PLAIN, hip.SOLVER_PERSISTENT):
        ctx = hip.Context(fullchem, 1024)
        ctx.set_solver(solver)
        conc = _t(T, w.conc)
        hnew = T.full((w.ncell,), 123.0, dtype=T.float64, device="cuda")
        hnew, ist, ierr, ok = ctx.fullchem_step(_t(T, w.met), _t(T, w.photol), _t(T, w.het), conc,
                                                _t(T, w.hstart), 1200.0, ic, box["rc"], hnew=hnew)
        assert not ok
        out[solver] = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy())
        ctx.close()
    a, b = out[hip.SOLVER_PLAIN], out[hip.SOLVER_PERSISTENT]
    assert (a[3] == -1).all() and (a[1] == 0.0).all() and (a[2] == 0).all()
    assert np.array_equal(a[0], np.maximum(w.conc, 0.0))
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_oversized_context_fails_cleanly(fullchem, torch_dev):
    """A context larger than the device fails with a message and leaves the
    device usable for the next context."""
    print("case: gckpp_init beyond HBM", flush=True)
    from gckpp_amd import hip
    with pytest.raises(hip.GckppError, match="cannot allocate"):
        hip.Context(fullchem, 2_000_000_000)
    ctx = hip.Context(fullchem, 64)
    ctx.close()


def test_handoff_and_split_settings_are_checked(fullchem, mech_of, torch_dev):
    """The tail hand-off and the split take a non-negative cell count, are
    refused for the carbon mechanism (no Rosenbrock solve), and report zero
    statistics before any launch used them."""
    print("case: hand-off and split argument checks", flush=True)
    from gckpp_amd import hip
    ctx = hip.Context(fullchem, 64)
    with pytest.raises(hip.GckppError, match="negative"):
        ctx.set_persistent_handoff(-1)
    with pytest.raises(hip.GckppError, match="negative"):
        ctx.set_persistent_split(-1)
    with pytest.raises(hip.GckppError, match="negative"):
        ctx.set_persistent_waves(-1)
    assert ctx.handoff_stats() == (0, 0, 0.0, 0.0)
    ctx.set_persistent_handoff(32)
    assert ctx.handoff_stats() == (0, 0, 0.0, 0.0)
    ctx.set_persistent_handoff(0)
    ctx.close()
    cc = hip.Context(mech_of("carbon"), 64)
    with pytest.raises(hip.GckppError, match="carbon"):
        cc.set_persistent_handoff(16)
    with pytest.raises(hip.GckppError, match="carbon"):
        cc.set_persistent_split(16)
    cc.close()
