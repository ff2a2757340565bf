"""GPU parity: the HIP kernels (through the C ABI) against the reference's
golden fixtures and the C oracle, on the same inputs.

Tolerances: Fun/Jac/LU/Solve are FP64 evaluations of the same sums in a
different order, so they agree to a few ulps of the largest term (1e-12
relative to the term magnitude).  Integrate takes data-dependent step-size
decisions, so results are compared with the reference's own scaled norm at
the repo tolerance (RTOL 0.5e-2, ATOL 1e-2): every species within
|c_gpu - c_ref| <= ATOL + RTOL*|c_ref|, i.e. scaled error <= 1.
"""
from pathlib import Path

import numpy as np
import pytest

from gckpp_amd import options

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


@pytest.fixture(scope="module")
def ctx(torch_dev, fullchem):
    from gckpp_amd import hip
    c = hip.Context(fullchem, 8192)
    yield c
    c.close()


@pytest.fixture(scope="module")
def gold():
    return dict(np.load(GOLD / "fullchem_box.npz"))


def _t(torch, a, dtype=None):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda", dtype=dtype or torch.float64)


def _term_mag_fun(m, conc, rct):
    """sum over reactions of |coef*A(r)| per species: the size of the terms
    whose sum Vdot is."""
    A = rct.copy()
    for r, rx in enumerate(m.reactions):
        for s, k in rx["reactants"]:
            A[r] *= conc[s] ** k
    mag = np.zeros((m.nvar, conc.shape[1]))
    for r, rx in enumerate(m.reactions):
        for s, c in rx["stoich"]:
            mag[s] += abs(c) * np.abs(A[r])
    return mag


def test_fun_matches_reference(ctx, gold, fullchem, torch_dev):
    print("case: Fun_SPLIT on golden box cells", flush=True)
    v = ctx.fun(_t(torch_dev, gold["conc"]), _t(torch_dev, gold["rconst"])).cpu().numpy()
    mag = _term_mag_fun(fullchem, gold["conc"], gold["rconst"])
    err = np.abs(v - gold["vdot"]) / (mag + 1e-300)
    assert err.max() < 1e-12, err.max()


def test_jac_matches_reference(ctx, gold, torch_dev):
    print("case: Jac_SP on golden box cells", flush=True)
    j = ctx.jac(_t(torch_dev, gold["conc"]), _t(torch_dev, gold["rconst"])).cpu().numpy()
    ref = gold["jvs"]
    scale = np.abs(ref).max(axis=0, keepdims=True)
    assert (np.abs(j - ref) / scale).max() < 1e-13
    # fill-in entries of the LU pattern are exact zeros in both
    assert np.array_equal(j == 0.0, ref == 0.0)


def test_decomp_and_solve_match_reference(ctx, gold, torch_dev):
    print("case: KppDecomp/KppSolve on golden box cells", flush=True)
    lu, ier = ctx.decomp(_t(torch_dev, gold["ghimj"]))
    lu = lu.cpu().numpy()
    assert np.array_equal(ier.cpu().numpy(), gold["ier"])
    ref = gold["lu"]
    assert (np.abs(lu - ref) / (np.abs(ref) + np.abs(ref).max(axis=0) * 1e-16)).max() < 1e-9
    xs = ctx.solve(_t(torch_dev, gold["lu"]), _t(torch_dev, gold["x"])).cpu().numpy()
    assert (np.abs(xs - gold["xs"]) / (np.abs(gold["xs"]).max(axis=0))).max() < 1e-12


def _scaled_err(c, ref, m, atol=1e-2, rtol=5e-3):
    nv = m.nvar
    return np.abs(c[:nv] - ref[:nv]) / (atol + rtol * np.abs(ref[:nv]))


def test_integrate_matches_reference(ctx, gold, fullchem, torch_dev):
    print("case: Integrate(0,1200) on golden box cells", flush=True)
    from gckpp_amd.options import fullchem_icntrl
    ic, rc = fullchem_icntrl()
    conc = _t(torch_dev, gold["conc"])
    ist, rst, ierr = ctx.integrate(conc, _t(torch_dev, gold["rconst"]),
                                   _t(torch_dev, np.zeros(conc.shape[1])), 0.0, 1200.0, ic, rc)
    c = conc.cpu().numpy()
    assert (ierr.cpu().numpy() == 1).all()
    assert _scaled_err(c, gold["int_conc"], fullchem).max() <= 1.0
    # same integrator decisions: step counts agree with the reference
    ist = ist.cpu().numpy()
    assert np.abs(ist[2] - gold["int_istatus"][2]).max() <= 2
    assert np.abs(rst.cpu().numpy()[0] - 1200.0).max() < 1e-9


def test_integrate_matches_oracle_on_box_workload(ctx, fullchem, torch_dev):
    """512 cells (256 dark, 256 lit) of the box workload against the C oracle
    with the reference's rates from the golden-checked oracle path."""
    print("case: Integrate 512 box cells vs C oracle", flush=True)
    from gckpp_amd import synth
    from oracle import orc, ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    idx = np.r_[0:256, 2048:2304]
    w = synth.box_cells(fullchem, 4096, 0).subset(idx)
    R = ref.RefKPP(fullchem)
    rct = R.rates(w)
    ic, rc = options.fullchem_icntrl()
    O = orc.Oracle(fullchem)
    c_o, i_o, r_o, e_o = O.integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc, nthreads=8)
    conc = _t(torch_dev, w.conc)
    ist, rst, ierr = ctx.integrate(conc, _t(torch_dev, rct), _t(torch_dev, w.hstart), 0.0, 1200.0, ic, rc)
    c = conc.cpu().numpy()
    assert (ierr.cpu().numpy() == e_o).all()
    err = _scaled_err(c, c_o, fullchem)
    print("max scaled err", err.max(), "median steps", np.median(ist.cpu().numpy()[2]))
    assert err.max() <= 1.0


@pytest.mark.parametrize("nblocks_cells", [512, 3000])
def test_persistent_solver_matches_oracle(fullchem, torch_dev, nblocks_cells):
    """The persistent work-stealing solver (row-Crout LU, per-lane slots)
    against the C oracle: same per-cell results at the repo tolerance, same
    step counts; more cells than resident lanes exercises the stealing."""
    print("case: persistent solver, %d box cells" % nblocks_cells, flush=True)
    from gckpp_amd import hip, synth
    from oracle import orc, ref
    if not ref.available():
        pytest.skip("oracle/_ref not built")
    w = synth.box_cells(fullchem, 4096, 0).subset(np.r_[0:nblocks_cells // 2, 2048:2048 + nblocks_cells // 2])
    R = ref.RefKPP(fullchem)
    rct = R.rates(w)
    ic, rc = options.fullchem_icntrl()
    O = orc.Oracle(fullchem)
    c_o, i_o, r_o, e_o = O.integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc, nthreads=8)
    ctx = hip.Context(fullchem, 4096)
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    conc = _t(torch_dev, w.conc)
    ist, rst, ierr = ctx.integrate(conc, _t(torch_dev, rct), _t(torch_dev, w.hstart), 0.0, 1200.0, ic, rc)
    c = conc.cpu().numpy()
    assert (ierr.cpu().numpy() == e_o).all()
    err = _scaled_err(c, c_o, fullchem)
    print("max scaled err", err.max())
    assert err.max() <= 1.0
    assert np.abs(ist.cpu().numpy()[2] - i_o[2]).max() <= 2
    assert np.abs(rst.cpu().numpy()[0] - 1200.0).max() < 1e-9
    ctx.close()
