"""Writes the golden fixtures of tests/golden/ from the compiled reference
(oracle/_ref, built by `make -C oracle ref` from the reference checkout).

Fixture <mech>_box.npz: a few box cells of the synthetic box workload (half
dark, half lit) with the reference's outputs on them:
  idx, met, photol, het, conc          inputs ([field][ncell])
  rconst                               Update_RCONST (inlined literals filled)
  vdot                                 Fun_SPLIT(conc, rconst)
  jvs                                  Jac_SP
  ghimj, lu, ier                       1/(h*gamma)-J with h=10 s, KppDecomp
  x, xs                                a random right-hand side, KppSolve
  int_conc, int_istatus, int_rstatus, int_ierr
                                       Integrate(0,1200) from Hstart=1e-5 s
  step_conc, step_hnew, step_ierr      Do_FullChem step (rates+integrate+retry+clip)
Hg_box.npz holds the same for the Hg mechanism on cells with rich per-cell
HetState records (its heterogeneous laws need cloud and aerosol), with the
Hg chemistry call's options (GeosCore/mercury_mod.F90:951-969).
carbon_box.npz (forward Euler, no Jacobian): inputs, rconst, vdot, the
Integrate result and step result, plus neg1_*/neg2_* for ICNTRL(16)=1 and 2
on rates scaled so that some results go negative.
Run from the repository root:  python tests/golden/make_golden.py [mech ...]
"""
import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(ROOT / "geos-chem_amd"))

from gckpp_amd import mechanism, options, synth  # noqa: E402
from oracle import ref  # noqa: E402


def make(mech_name, cells, het_seed=None):
    m = mechanism.load(mech_name)
    R = ref.RefKPP(m)
    w = synth.box_cells(m, 4096, 0).subset(cells)
    if het_seed is not None:
        w.het = synth.rich_het(np.random.default_rng(het_seed), len(cells))
    ic, rc = options.icntrl_for(mech_name)
    rct = R.rates(w)
    vdot = R.fun(w.conc, rct)
    jvs = R.jac(w.conc, rct)
    ghimj = -jvs.copy()
    ghimj[m.lu_diag[: m.nvar]] += 1.0 / (10.0 * 0.5)
    lu, ier = R.decomp(ghimj)
    x = np.random.default_rng(7).standard_normal((m.nvar, len(cells)))
    xs = R.solve(lu, x)
    c1, i1, r1, e1 = R.integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc)
    c2, i2, h2, e2 = R.fullchem_step(w, ic, rc, nthreads=1)
    out = ROOT / "tests" / "golden" / ("%s_box.npz" % mech_name)
    np.savez_compressed(out, idx=np.asarray(cells), met=w.met, photol=w.photol, het=w.het,
                        conc=w.conc, rconst=rct, vdot=vdot, jvs=jvs, ghimj=ghimj, lu=lu, ier=ier,
                        x=x, xs=xs, int_conc=c1, int_istatus=i1, int_rstatus=r1, int_ierr=e1,
                        step_conc=c2, step_hnew=h2, step_ierr=e2)
    print("wrote", out, out.stat().st_size, "bytes; steps", i1[2])


def make_het(mech_name, ncell=64, seed=11):
    """Rates of box cells with rich per-cell HetState records (clouds, ice,
    every aerosol type, random pH and halides): pins the heterogeneous laws."""
    m = mechanism.load(mech_name)
    R = ref.RefKPP(m)
    w = synth.box_cells(m, 4096, 0).subset(np.r_[0:ncell // 2, 2048:2048 + ncell // 2])
    w.het = synth.rich_het(np.random.default_rng(seed), ncell)
    rct = R.rates(w)
    out = ROOT / "tests" / "golden" / ("%s_het.npz" % mech_name)
    np.savez_compressed(out, met=w.met, photol=w.photol, het=w.het, conc=w.conc, rconst=rct)
    print("wrote", out, out.stat().st_size, "bytes; nonzero het-driven rates",
          int((rct[780:900] != 0).sum()))


def make_carbon(cells):
    m = mechanism.load("carbon")
    R = ref.RefKPP(m)
    w = synth.box_cells(m, 4096, 0).subset(cells)
    ic, rc = options.carbon_icntrl()
    rct = R.rates(w)
    vdot = R.fun(w.conc, rct)
    c1, _, _, e1 = R.integrate(w.conc, rct, w.hstart, 0.0, 1200.0, ic, rc)
    c2, _, _, e2 = R.fullchem_step(w, ic, rc, nthreads=1)
    # strong first-order losses (CH4, CO in the stratosphere; CO+OH) make
    # some Euler results negative
    rneg = rct.copy()
    rneg[[3, 5, 7]] *= 1.0e4
    extra = {}
    for mode in (1, 2):
        icm = ic.copy()
        icm[15] = mode
        cn, _, _, en = R.integrate(w.conc, rneg, w.hstart, 0.0, 1200.0, icm, rc)
        extra["neg%d_conc" % mode], extra["neg%d_ierr" % mode] = cn, en
    out = ROOT / "tests" / "golden" / "carbon_box.npz"
    np.savez_compressed(out, idx=np.asarray(cells), met=w.met, photol=w.photol, het=w.het,
                        conc=w.conc, rconst=rct, vdot=vdot, int_conc=c1, int_ierr=e1,
                        step_conc=c2, step_ierr=e2, rneg=rneg, **extra)
    print("wrote", out, out.stat().st_size, "bytes; mode-2 failures",
          int((extra["neg2_ierr"] < 0).sum()))


if __name__ == "__main__":
    todo = sys.argv[1:] or ["fullchem", "Hg", "carbon"]
    if "fullchem" in todo:
        make("fullchem", [0, 1, 2, 3, 2048, 2049, 2050, 2051])
        make_het("fullchem")
    if "Hg" in todo:
        make("Hg", list(range(0, 16)) + list(range(2048, 2064)), het_seed=5)
    if "carbon" in todo:
        make_carbon(list(range(0, 32)) + list(range(2048, 2080)))
