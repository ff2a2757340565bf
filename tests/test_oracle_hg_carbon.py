"""CPU checks that pin the C oracle on the Hg and carbon mechanisms.

Hg (Rodas3, KPP/Hg): Fun, Jac, LU and Solve against the reference's own
outputs (tests/golden/Hg_box.npz, written from the compiled reference) and
against numpy (mass action, dense LU, numpy.linalg.solve); Integrate with the
Hg chemistry call's options against the reference's ISTATUS and results.

carbon (forward Euler, KPP/carbon/gckpp_Integrator.F90:155-215): Fun and the
Euler step against the reference's outputs (carbon_box.npz), including the
ICNTRL(16) negative-value modes, and against the closed form y0 + f(y0)*dt.
"""
from pathlib import Path

import numpy as np
import pytest

GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def carbon():
    from gckpp_amd import mechanism
    return mechanism.load("carbon")


def _oracle(m):
    from oracle import orc
    if not orc.lib_path().exists():
        pytest.skip("oracle not built (make -C oracle oracle)")
    return orc.Oracle(m)


def _mass_action(m, conc, rct):
    """Vdot and the magnitude of its terms, straight from the reaction list."""
    A = rct.copy()
    for r, rx in enumerate(m.reactions):
        for s, k in rx["reactants"]:
            A[r] *= conc[s] ** k
    vd = np.zeros((m.nvar, conc.shape[1]))
    mag = np.zeros_like(vd)
    for r, rx in enumerate(m.reactions):
        for s, c in rx["stoich"]:
            vd[s] += c * A[r]
            mag[s] += abs(c) * np.abs(A[r])
    return vd, mag


def _dense(m, vals):
    A = np.zeros((m.nvar, m.nvar))
    A[m.lu_irow, m.lu_icol] = vals
    return A


# ----------------------------------------------------------------------- Hg
@pytest.fixture(scope="module")
def hg_gold():
    return dict(np.load(GOLD / "Hg_box.npz"))


def test_hg_fun_jac_match_reference(hg, hg_gold):
    O = _oracle(hg)
    g = hg_gold
    v = O.fun(g["conc"], g["rconst"])
    vd, mag = _mass_action(hg, g["conc"], g["rconst"])
    assert (np.abs(v - g["vdot"]) / (mag + 1e-300)).max() < 1e-13
    assert (np.abs(v - vd) / (mag + 1e-300)).max() < 1e-13
    j = O.jac(g["conc"], g["rconst"])
    assert np.array_equal(j, g["jvs"])


def test_hg_decomp_solve_match_reference_and_numpy(hg, hg_gold):
    O = _oracle(hg)
    g = hg_gold
    lu, ier = O.decomp(g["ghimj"])
    assert np.array_equal(ier, g["ier"])
    assert np.array_equal(lu, g["lu"])
    xs = O.solve(g["lu"], g["x"])
    assert np.array_equal(xs, g["xs"])
    for col in (0, 17):
        A = _dense(hg, g["ghimj"][:, col])
        ref = np.linalg.solve(A, g["x"][:, col])
        assert np.abs(xs[:, col] - ref).max() / np.abs(ref).max() < 1e-10


def test_hg_integrate_matches_reference(hg, hg_gold):
    from gckpp_amd.options import hg_icntrl
    O = _oracle(hg)
    g = hg_gold
    ic, rc = hg_icntrl()
    n = g["conc"].shape[1]
    c, ist, rst, ierr = O.integrate(g["conc"], g["rconst"], np.zeros(n), 0.0, 1200.0, ic, rc)
    assert np.array_equal(ierr, g["int_ierr"])
    assert np.array_equal(ist[:8], g["int_istatus"][:8])
    nv = hg.nvar
    ref = g["int_conc"]
    err = np.abs(c[:nv] - ref[:nv]) / (hg.atol()[:, None] + hg.rtol()[:, None] * np.abs(ref[:nv]))
    assert err.max() < 1e-3


def test_hg_het_rates_are_exercised(hg, hg_gold):
    """The fixture's cells reach the Hg heterogeneous laws (cloud boxes with
    liquid and ice), so rate parity covers them."""
    het_rx = [i for i, r in enumerate(hg.reactions) if "Het_Hg" in r["rate"]]
    assert len(het_rx) == 35
    assert (hg_gold["rconst"][het_rx] > 0).sum() > 100


# ------------------------------------------------------------------- carbon
@pytest.fixture(scope="module")
def carbon_gold():
    return dict(np.load(GOLD / "carbon_box.npz"))


def _close(m, c, ref, conc0, rct, dt=1200.0):
    """Euler results agree to roundoff of y0 + dt*sum(terms): the reference's
    Fun adds the same terms in another order."""
    _, mag = _mass_action(m, conc0, rct)
    scale = np.abs(conc0[:m.nvar]) + dt * mag
    assert (np.abs(c[:m.nvar] - ref[:m.nvar]) <= 1e-14 * scale).all()
    assert np.array_equal(c[m.nvar:], ref[m.nvar:])
    assert np.array_equal(c[:m.nvar] == 0.0, ref[:m.nvar] == 0.0)


def test_carbon_fun_matches_reference(carbon, carbon_gold):
    O = _oracle(carbon)
    g = carbon_gold
    v = O.fun(g["conc"], g["rconst"])
    vd, mag = _mass_action(carbon, g["conc"], g["rconst"])
    assert (np.abs(v - g["vdot"]) / (mag + 1e-300)).max() < 1e-14
    assert (np.abs(v - vd) / (mag + 1e-300)).max() < 1e-14


def test_carbon_euler_matches_reference_and_closed_form(carbon, carbon_gold):
    from gckpp_amd.options import carbon_icntrl
    O = _oracle(carbon)
    g = carbon_gold
    ic, _ = carbon_icntrl()
    c, ierr = O.euler(g["conc"], g["rconst"], 0.0, 1200.0, ic)
    _close(carbon, c, g["int_conc"], g["conc"], g["rconst"])
    assert (ierr == 1).all() and np.array_equal(ierr, g["int_ierr"])
    vd, _ = _mass_action(carbon, g["conc"], g["rconst"])
    nv = carbon.nvar
    expect = g["conc"][:nv] + vd * 1200.0
    assert np.allclose(c[:nv], expect, rtol=1e-13, atol=1e-300)
    # fixed species are not touched
    assert np.array_equal(c[nv:], g["conc"][nv:])


def test_carbon_negative_modes_match_reference(carbon, carbon_gold):
    from gckpp_amd.options import carbon_icntrl
    O = _oracle(carbon)
    g = carbon_gold
    ic, _ = carbon_icntrl()
    c0, _ = O.euler(g["conc"], g["rneg"], 0.0, 1200.0, ic)
    assert (c0[:carbon.nvar] < 0).any(), "the scaled rates must drive some species negative"
    for mode in (1, 2):
        icm = ic.copy()
        icm[15] = mode
        c, ierr = O.euler(g["conc"], g["rneg"], 0.0, 1200.0, icm)
        assert np.array_equal(ierr, g["neg%d_ierr" % mode])
        _close(carbon, c, g["neg%d_conc" % mode], g["conc"], g["rneg"])
    # mode 1 clips to zero, mode 2 leaves failing cells unchanged
    assert (g["neg1_conc"][:carbon.nvar] >= 0).all()
    bad = g["neg2_ierr"] == -9
    assert bad.any() and np.array_equal(g["neg2_conc"][:, bad], g["conc"][:, bad])
