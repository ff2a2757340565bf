"""The mechanism compiler from KPP source (geos-chem_amd/mechgen/eqn_compile.py),
on the CPU, against the reference's KPP sources and KPP's own generated code.

  * With KPP's species order (gckpp_Parameters.F90), compiling
    KPP/<mech>/<mech>.kpp (which includes <mech>.eqn) gives exactly the
    tables read out of KPP's generated Fortran (mechgen/kpp_parse.py) and
    the committed mech/<mech>.json the build wrote: species, every reaction's
    reactants, net stoichiometry, rate expression and inlining, the
    Jacobian's nonzero count and the LU pattern -- for fullchem, Hg and
    carbon.
  * With the compiler's own Markowitz order, the LU pattern is a closed
    symbolic factorization of the Jacobian pattern (a dense no-pivot LU of
    a random matrix on it puts nothing outside it) with no more fill than
    KPP's, and the C oracle on the reordered fullchem reproduces the
    compiled reference's Fun and Integrate on box cells (species permuted).
The reference checkout is read where it is (the tests skip without it).
"""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
REF = Path("/root/reference")
sys.path.insert(0, str(ROOT / "geos-chem_amd" / "mechgen"))


def _src(mech):
    kpp = REF / "KPP" / mech / ("%s.kpp" % mech)
    if not kpp.exists():
        pytest.skip("reference KPP sources not present")
    return kpp, REF / "KPP" / mech / "gckpp_Parameters.F90"


@pytest.mark.parametrize("name", ["fullchem", "Hg", "carbon"])
def test_kpp_order_reproduces_kpp_tables(name, mech_of):
    """Against KPP's generated Fortran, read by kpp_parse.py, and against the
    committed mechanism the build generated."""
    import eqn_compile as E
    import kpp_parse
    kpp, par = _src(name)
    d = E.compile_mechanism(kpp, name, par)
    ref = kpp_parse.build(REF, name)
    committed = mech_of(name).d
    for k in ("species", "reactions", "lu_irow", "lu_icol", "lu_crow", "lu_diag", "nonzero"):
        assert committed[k] == d[k], k
    for k in ("nspec", "nvar", "nfix", "nreact", "nonzero", "lu_nonzero", "species",
              "lu_irow", "lu_icol", "lu_crow", "lu_diag"):
        assert d[k] == ref[k], k
    assert len(d["reactions"]) == len(ref["reactions"])
    for r, (a, b) in enumerate(zip(d["reactions"], ref["reactions"])):
        assert a["reactants"] == b["reactants"], r
        assert a["stoich"] == b["stoich"], r
        assert a["rate"] == b["rate"], r
        assert a["inlined"] == b["inlined"], r


def _dense_pattern(d):
    nv = d["nvar"]
    P = np.zeros((nv, nv), dtype=bool)
    P[np.asarray(d["lu_irow"]), np.asarray(d["lu_icol"])] = True
    return P


def test_markowitz_order_is_a_closed_factorization_with_less_fill(fullchem):
    import eqn_compile as E
    kpp, _ = _src("fullchem")
    d = E.compile_mechanism(kpp, "fullchem")
    assert sorted(d["species"]) == sorted(fullchem.species)
    assert d["species"][d["nvar"]:] == fullchem.species[fullchem.nvar:]   # fixed species last
    assert d["nonzero"] == fullchem.d["nonzero"]
    assert d["lu_nonzero"] <= fullchem.lu_nonzero
    nv = d["nvar"]
    P = _dense_pattern(d)
    # the Jacobian pattern is inside, and every diagonal entry is present
    idx = {n: i for i, n in enumerate(d["species"])}
    J = np.eye(nv, dtype=bool)
    for rx in d["reactions"]:
        cols = [s for s, _ in rx["reactants"] if s < nv]
        for s, _ in rx["stoich"]:
            J[s, cols] = True
    assert (P | ~J).all()
    assert len(idx) == d["nspec"]
    # a dense no-pivot LU of a random diagonally dominant matrix on the
    # Jacobian pattern has no nonzero outside the LU pattern
    rng = np.random.default_rng(1)
    A = np.where(J, rng.uniform(-1, 1, (nv, nv)), 0.0) + np.eye(nv) * nv
    U = A.copy()
    L = np.eye(nv)
    for k in range(nv):
        f = U[k + 1:, k] / U[k, k]
        L[k + 1:, k] = f
        U[k + 1:, :] -= np.outer(f, U[k, :])
    F = np.tril(L, -1) + np.triu(U)
    assert np.count_nonzero(F[~P]) == 0
    # row-wise storage: columns ascending, diagonal pointer on the diagonal
    ic, cr, dg = d["lu_icol"], d["lu_crow"], d["lu_diag"]
    for i in range(nv):
        row = ic[cr[i]:cr[i + 1]]
        assert row == sorted(row) and ic[dg[i]] == i


def test_markowitz_fullchem_matches_the_reference_solve(fullchem):
    """The C oracle on the reordered mechanism (its own LU order) against the
    compiled reference on the same cells and rate constants."""
    import eqn_compile as E
    from gckpp_amd import mechanism, options, synth
    from oracle import orc, ref
    kpp, _ = _src("fullchem")
    if not orc.lib_path().exists() or not ref.available("fullchem"):
        pytest.skip("oracles not built")
    d = E.compile_mechanism(kpp, "fullchem")
    m2 = mechanism.reordered(fullchem, d)
    pos = {n: i for i, n in enumerate(fullchem.species)}
    perm = np.array([pos[n] for n in d["species"]])
    w = synth.box_cells(fullchem, 4096, 0).subset(np.r_[0:24, 2048:2072])
    R = ref.RefKPP(fullchem)
    rct = R.rates(w)
    ic, rc = options.fullchem_icntrl()
    O2 = orc.Oracle(m2)
    c2 = np.ascontiguousarray(w.conc[perm])
    # Fun: the same rates of change, permuted
    v_ref = R.fun(w.conc, rct)
    v2 = O2.fun(c2, rct)
    nv = fullchem.nvar
    inv = np.empty(nv, dtype=np.int64)
    inv[perm[:nv]] = np.arange(nv)
    mag = np.abs(v_ref).max(axis=0) + 1e-300
    assert (np.abs(v2[inv] - v_ref) / mag).max() < 1e-12
    # Integrate: the repo tolerance, the same step counts within 2
    c_r, ist_r, _, e_r = R.integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc)
    c_o, ist_o, _, e_o = O2.integrate(c2, rct, w.hstart, 0.0, 1200.0, ic, rc)
    assert (e_r == 1).all() and (e_o == 1).all()
    co = c_o[inv]
    sc = np.abs(co - c_r[:nv]) / (fullchem.atol()[:, None] + fullchem.rtol()[:, None] * np.abs(c_r[:nv]))
    assert sc.max() <= 1.0
    assert np.abs(ist_o[2].astype(int) - ist_r[2]).max() <= 2
