"""The step-size error branches on the GPU against the compiled reference
(oracle/_ref) on the same inputs and the same rate constants: IERR -7 at
the first check and a few steps in, and a singular first decomposition that
halves the step (ISTATUS(8)), in the lane-per-cell solver, the persistent
solver and its cell-per-wave variant.  The inputs are those of
tests/error_cases.py.  Reference: KPP/fullchem/gckpp_Integrator.F90:659-662
(-7), :686-690 and :1921-1999 (PrepareMatrix), gckpp_LinearAlgebra.F90:61-65.
"""
import numpy as np
import pytest

import error_cases as E

pytestmark = pytest.mark.gpu

CONFIGS = [("plain", None), ("persistent", None), ("persistent", "wave")]


@pytest.fixture(scope="module")
def case(fullchem):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from gckpp_amd import options, synth
    from oracle import ref
    if not ref.available("fullchem"):
        pytest.skip("oracle/_ref not built")
    w = synth.box_cells(fullchem, 4096, 0).subset(np.r_[0:256, 2048:2304])
    R = ref.RefKPP(fullchem)
    ic, rc = options.fullchem_icntrl()
    return dict(T=torch, w=w, R=R, rct=R.rates(w), ic=ic, rc=rc)


def _gpu(case, m, conc, rct, hstart, tin, tout, solver, variant):
    from gckpp_amd import hip
    T = case["T"]
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    ctx = hip.Context(m, conc.shape[1])
    ctx.set_solver(hip.SOLVER_PLAIN if solver == "plain" else hip.SOLVER_PERSISTENT)
    if variant == "wave":
        ctx.set_persistent_config(hip.WAVE_VARIANT)
    c = dev(conc)
    ist, rst, ierr = ctx.integrate(c, dev(rct), dev(hstart), tin, tout, case["ic"], case["rc"])
    T.cuda.synchronize()
    out = c.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy()
    ctx.close()
    return out


def _compare(m, label, gpu, ref):
    c, ist, ierr = gpu
    c_r, ist_r, e_r = ref
    nv = m.nvar
    sc = np.abs(c[:nv] - c_r[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(c_r[:nv]))
    print("  %s: IERR %s, max scaled error vs reference %.3g" %
          (label, dict(zip(*[x.tolist() for x in np.unique(ierr, return_counts=True)])), sc.max()),
          flush=True)
    assert np.array_equal(ierr, e_r), label
    assert np.array_equal(ist[:8], ist_r[:8]), label
    assert sc.max() <= 1e-3, label


@pytest.mark.parametrize("solver,variant", CONFIGS)
def test_step_too_small_at_the_first_check(solver, variant, case, fullchem):
    print("case: -7 at the first check, %s %s" % (solver, variant or ""), flush=True)
    w, t0 = case["w"], E.T_TINY_FIRST
    c, ist, ierr = _gpu(case, fullchem, w.conc, case["rct"], w.hstart, t0, t0 + 1200.0, solver, variant)
    assert (ierr == -7).all()
    assert (ist == 0).all()
    assert np.array_equal(c, w.conc)


@pytest.mark.parametrize("solver,variant", CONFIGS)
def test_step_too_small_after_some_steps(solver, variant, case, fullchem):
    print("case: -7 a few steps in, %s %s" % (solver, variant or ""), flush=True)
    w, t0 = case["w"], E.T_TINY_LATER
    h = np.full(w.ncell, E.H_LATER)
    c_r, ist_r, _, e_r = case["R"].integrate(w.conc, case["rct"], h, t0, t0 + 1200.0, case["ic"], case["rc"])
    assert 0 < (e_r == -7).sum() < w.ncell // 2
    gpu = _gpu(case, fullchem, w.conc, case["rct"], h, t0, t0 + 1200.0, solver, variant)
    _compare(fullchem, "-7 later", gpu, (c_r, ist_r, e_r))


@pytest.mark.parametrize("solver,variant", CONFIGS)
def test_singular_decomposition_halves_the_step(solver, variant, case, fullchem):
    print("case: singular first decomposition, %s %s" % (solver, variant or ""), flush=True)
    w = case["w"]
    conc, rct, h, tout = E.singular_cells(fullchem, w, case["rct"])
    c_r, ist_r, _, e_r = case["R"].integrate(conc, rct, h, 0.0, tout, case["ic"], case["rc"])
    assert np.array_equal(ist_r[7], (np.arange(w.ncell) % 2 == 0).astype(np.int32))
    gpu = _gpu(case, fullchem, conc, rct, h, 0.0, tout, solver, variant)
    _compare(fullchem, "singular", gpu, (c_r, ist_r, e_r))
