"""GPU parity of the Hg (Rodas3) and carbon (forward Euler) mechanisms
through the C ABI, against the reference's golden outputs
(tests/golden/Hg_box.npz, carbon_box.npz, written from the compiled
reference).

Tolerances: rate programs and Fun/Jac/LU/Solve are FP64 evaluations of the
same expressions, possibly contracted to FMAs or summed in another order:
1e-12 relative (Jac/LU/Solve: relative to the column scale).  Integrate and
the step are compared with the reference's scaled norm at the mechanism's
tolerances (Hg: RTOL 1e-2 default, ATOL 1e-2): scaled error <= 1.  The
carbon Euler step is one evaluation: within 1e-13 of y0 + dt*sum|terms|.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = Path(__file__).resolve().parent / "golden"


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


def _rel_rates(out, ref):
    assert np.array_equal(out == 0.0, ref == 0.0)
    nz = ref != 0
    return (np.abs(out[nz] - ref[nz]) / np.abs(ref[nz])).max()


def _term_mag(m, conc, rct):
    A = rct.copy()
    for r, rx in enumerate(m.reactions):
        for s, k in rx["reactants"]:
            A[r] *= conc[s] ** k
    mag = np.zeros((m.nvar, conc.shape[1]))
    for r, rx in enumerate(m.reactions):
        for s, c in rx["stoich"]:
            mag[s] += abs(c) * np.abs(A[r])
    return mag


# ----------------------------------------------------------------------- Hg
@pytest.fixture(scope="module")
def hg_gold():
    return dict(np.load(GOLD / "Hg_box.npz"))


@pytest.fixture(scope="module")
def hg_ctx(torch_dev, hg):
    from gckpp_amd import hip
    c = hip.Context(hg, 1024)
    yield c
    c.close()


def test_hg_rates_match_reference(hg_ctx, hg_gold, torch_dev):
    print("case: Hg Update_RCONST, 32 box cells with per-cell HetState", flush=True)
    g = hg_gold
    out = hg_ctx.rates(_t(torch_dev, g["met"]), _t(torch_dev, g["photol"]), _t(torch_dev, g["het"]),
                       _t(torch_dev, g["conc"])).cpu().numpy()
    assert _rel_rates(out, g["rconst"]) < 1e-12


def test_hg_components_match_reference(hg_ctx, hg_gold, hg, torch_dev):
    print("case: Hg Fun/Jac/Decomp/Solve on golden cells", flush=True)
    g = hg_gold
    conc, rct = _t(torch_dev, g["conc"]), _t(torch_dev, g["rconst"])
    v = hg_ctx.fun(conc, rct).cpu().numpy()
    assert (np.abs(v - g["vdot"]) / (_term_mag(hg, g["conc"], g["rconst"]) + 1e-300)).max() < 1e-12
    j = hg_ctx.jac(conc, rct).cpu().numpy()
    scale = np.abs(g["jvs"]).max(axis=0, keepdims=True)
    assert (np.abs(j - g["jvs"]) / scale).max() < 1e-13
    lu, ier = hg_ctx.decomp(_t(torch_dev, g["ghimj"]))
    assert np.array_equal(ier.cpu().numpy(), g["ier"])
    lu = lu.cpu().numpy()
    assert (np.abs(lu - g["lu"]) / (np.abs(g["lu"]) + np.abs(g["lu"]).max(axis=0) * 1e-16)).max() < 1e-9
    xs = hg_ctx.solve(_t(torch_dev, g["lu"]), _t(torch_dev, g["x"])).cpu().numpy()
    assert (np.abs(xs - g["xs"]) / np.abs(g["xs"]).max(axis=0)).max() < 1e-12


def _scaled(m, c, ref):
    nv = m.nvar
    return np.abs(c[:nv] - ref[:nv]) / (m.atol()[:, None] + m.rtol()[:, None] * np.abs(ref[:nv]))


@pytest.mark.parametrize("solver", ["plain", "persistent"])
def test_hg_integrate_matches_reference(hg, hg_gold, torch_dev, solver):
    print("case: Hg Integrate(0,1200), %s solver" % solver, flush=True)
    from gckpp_amd import hip
    from gckpp_amd.options import hg_icntrl
    g = hg_gold
    ctx = hip.Context(hg, 1024)
    if solver == "persistent":
        ctx.set_solver(hip.SOLVER_PERSISTENT)
    ic, rc = hg_icntrl()
    conc = _t(torch_dev, g["conc"])
    n = conc.shape[1]
    ist, rst, ierr = ctx.integrate(conc, _t(torch_dev, g["rconst"]), _t(torch_dev, np.zeros(n)),
                                   0.0, 1200.0, ic, rc)
    c = conc.cpu().numpy()
    assert np.array_equal(ierr.cpu().numpy(), g["int_ierr"])
    assert _scaled(hg, c, g["int_conc"]).max() <= 1.0
    assert np.abs(ist.cpu().numpy()[2] - g["int_istatus"][2]).max() <= 2
    assert np.abs(rst.cpu().numpy()[0] - 1200.0).max() < 1e-9
    ctx.close()


def test_hg_step_matches_reference(hg_ctx, hg_gold, hg, torch_dev):
    """The chemistry step of the Hg module (rates, Integrate, retry, clip),
    mercury_mod.F90:1033-1260, against the reference's."""
    print("case: Hg chemistry step on golden cells", flush=True)
    from gckpp_amd.options import hg_icntrl
    g = hg_gold
    ic, rc = hg_icntrl()
    conc = _t(torch_dev, g["conc"])
    n = conc.shape[1]
    hnew, ist, ierr, ok = hg_ctx.fullchem_step(_t(torch_dev, g["met"]), _t(torch_dev, g["photol"]),
                                               _t(torch_dev, g["het"]), conc,
                                               _t(torch_dev, np.zeros(n)), 1200.0, ic, rc)
    assert ok
    assert np.array_equal(ierr.cpu().numpy(), g["step_ierr"])
    assert _scaled(hg, conc.cpu().numpy(), g["step_conc"]).max() <= 1.0
    h = hnew.cpu().numpy()
    assert np.all(np.abs(h - g["step_hnew"]) <= 0.5 * g["step_hnew"])


# ------------------------------------------------------------------- carbon
@pytest.fixture(scope="module")
def carbon_gold():
    return dict(np.load(GOLD / "carbon_box.npz"))


@pytest.fixture(scope="module")
def carbon_ctx(torch_dev, carbon):
    from gckpp_amd import hip
    c = hip.Context(carbon, 1024)
    yield c
    c.close()


def _euler_close(m, c, ref, conc0, rct, dt=1200.0):
    nv = m.nvar
    scale = np.abs(conc0[:nv]) + dt * _term_mag(m, conc0, rct)
    assert (np.abs(c[:nv] - ref[:nv]) <= 1e-13 * scale).all()
    assert np.array_equal(c[nv:], ref[nv:])
    assert np.array_equal(c[:nv] == 0.0, ref[:nv] == 0.0)


def test_carbon_rates_and_fun_match_reference(carbon_ctx, carbon_gold, carbon, torch_dev):
    print("case: carbon Update_RCONST and Fun, 64 box cells", flush=True)
    g = carbon_gold
    conc = _t(torch_dev, g["conc"])
    out = carbon_ctx.rates(_t(torch_dev, g["met"]), _t(torch_dev, g["photol"]), _t(torch_dev, g["het"]),
                           conc).cpu().numpy()
    assert _rel_rates(out, g["rconst"]) < 1e-12
    v = carbon_ctx.fun(conc, _t(torch_dev, g["rconst"])).cpu().numpy()
    mag = _term_mag(carbon, g["conc"], g["rconst"])
    assert (np.abs(v - g["vdot"]) / (mag + 1e-300)).max() < 1e-13


def test_carbon_euler_matches_reference(carbon_ctx, carbon_gold, carbon, torch_dev):
    print("case: carbon forward Euler, ICNTRL(16) = 0, 1, 2", flush=True)
    from gckpp_amd.options import carbon_icntrl
    g = carbon_gold
    ic, rc = carbon_icntrl()
    n = g["conc"].shape[1]
    h0 = _t(torch_dev, np.zeros(n))
    for mode, rct, cref, eref in ((0, g["rconst"], g["int_conc"], g["int_ierr"]),
                                  (1, g["rneg"], g["neg1_conc"], g["neg1_ierr"]),
                                  (2, g["rneg"], g["neg2_conc"], g["neg2_ierr"])):
        icm = ic.copy()
        icm[15] = mode
        conc = _t(torch_dev, g["conc"])
        _, _, ierr = carbon_ctx.integrate(conc, _t(torch_dev, rct), h0, 0.0, 1200.0, icm, rc)
        assert np.array_equal(ierr.cpu().numpy(), eref), mode
        _euler_close(carbon, conc.cpu().numpy(), cref, g["conc"], rct)


def test_carbon_step_matches_reference(carbon_ctx, carbon_gold, carbon, torch_dev):
    """The carbon chemistry step (carbon_gases_mod.F90:450-604): rates, one
    Euler step, no clipping of negatives."""
    print("case: carbon chemistry step", flush=True)
    from gckpp_amd.options import carbon_icntrl
    g = carbon_gold
    ic, rc = carbon_icntrl()
    conc = _t(torch_dev, g["conc"])
    n = conc.shape[1]
    _, _, ierr, ok = carbon_ctx.fullchem_step(_t(torch_dev, g["met"]), _t(torch_dev, g["photol"]),
                                              _t(torch_dev, g["het"]), conc,
                                              _t(torch_dev, np.zeros(n)), 1200.0, ic, rc)
    assert ok and np.array_equal(ierr.cpu().numpy(), g["step_ierr"])
    _euler_close(carbon, conc.cpu().numpy(), g["step_conc"], g["conc"], g["rconst"])


def test_carbon_has_no_jacobian(carbon_ctx, carbon_gold, torch_dev):
    from gckpp_amd import hip
    g = carbon_gold
    with pytest.raises(hip.GckppError):
        carbon_ctx.jac(_t(torch_dev, g["conc"]), _t(torch_dev, g["rconst"]))
