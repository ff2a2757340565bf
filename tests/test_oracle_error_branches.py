"""The C oracle's step-size error branches against the compiled reference
(oracle/_ref) and against what the arithmetic fixes by itself.

  -7 at the first check (TIN = 1e12 s): 1e12 + 1e-6 == 1e12 in doubles, so
     every cell returns -7 with no work done (ISTATUS all zero) and its
     concentrations untouched;
  -7 a few steps in (TIN = 2**42 s, Hstart = 1 s): the same cells stop,
     after the same step counts, as in the reference;
  singular first decomposition (CH2I2 photolysis rate -2048 /s, Hstart
     2**-10 s): the Ghimj diagonal of CH2I2 is exactly zero, so exactly the
     cells with that rate report one singular decomposition (ISTATUS(8)).
Reference: KPP/fullchem/gckpp_Integrator.F90:659-662 (-7), :686-690 and
:1921-1999 (PrepareMatrix, -8), gckpp_LinearAlgebra.F90:61-65 (pivot check).
"""
import numpy as np
import pytest

import error_cases as E


@pytest.fixture(scope="module")
def case(fullchem):
    from gckpp_amd import options, synth
    from oracle import orc, ref
    if not orc.lib_path().exists():
        pytest.skip("oracle not built")
    if not ref.available("fullchem"):
        pytest.skip("oracle/_ref not built")
    w = synth.box_cells(fullchem, 4096, 0).subset(np.r_[0:32, 2048:2080])
    R = ref.RefKPP(fullchem)
    ic, rc = options.fullchem_icntrl()
    return dict(w=w, R=R, O=orc.Oracle(fullchem), rct=R.rates(w), ic=ic, rc=rc)


def _scaled(m, c, ref):
    nv = m.nvar
    return (np.abs(c[:nv] - ref[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(ref[:nv]))).max()


def test_step_too_small_at_the_first_check(case, fullchem):
    w = case["w"]
    t0 = E.T_TINY_FIRST
    assert t0 + 0.1 * 1e-5 == t0
    for run in (case["R"].integrate, case["O"].integrate):
        c, ist, rst, ierr = run(w.conc, case["rct"], w.hstart, t0, t0 + 1200.0, case["ic"], case["rc"])
        assert (ierr == -7).all()
        assert (ist == 0).all()
        assert np.array_equal(c, w.conc)


def test_step_too_small_after_some_steps(case, fullchem):
    w = case["w"]
    t0 = E.T_TINY_LATER
    h = np.full(w.ncell, E.H_LATER)
    c_r, ist_r, _, e_r = case["R"].integrate(w.conc, case["rct"], h, t0, t0 + 1200.0, case["ic"], case["rc"])
    c_o, ist_o, _, e_o = case["O"].integrate(w.conc, case["rct"], h, t0, t0 + 1200.0, case["ic"], case["rc"])
    stopped = e_r == -7
    assert 0 < stopped.sum() < w.ncell // 2
    assert (ist_r[2][stopped] > 0).all()          # after accepted steps
    assert np.array_equal(e_o, e_r)
    assert np.array_equal(ist_o[:8], ist_r[:8])
    assert _scaled(fullchem, c_o, c_r) < 1e-3


def test_singular_decomposition_halves_the_step(case, fullchem):
    w = case["w"]
    conc, rct, h, tout = E.singular_cells(fullchem, w, case["rct"])
    c_r, ist_r, _, e_r = case["R"].integrate(conc, rct, h, 0.0, tout, case["ic"], case["rc"])
    c_o, ist_o, _, e_o = case["O"].integrate(conc, rct, h, 0.0, tout, case["ic"], case["rc"])
    want = (np.arange(w.ncell) % 2 == 0).astype(np.int32)
    assert np.array_equal(ist_r[7], want)
    assert (e_r == 1).all()
    assert np.array_equal(e_o, e_r)
    assert np.array_equal(ist_o[:8], ist_r[:8])
    assert _scaled(fullchem, c_o, c_r) < 1e-3
