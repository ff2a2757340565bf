"""The cell-per-wave solver's programs (mechgen/wave_sched.py) perform every
sum of the solve with the same terms in the same order as the lane-per-cell
solvers' table walk: executed on the CPU (one IEEE double operation per
step, as a lane does it) they give bit-identical Vdot, Jacobian, LU and
substitution results on random data with the mechanism's patterns, and every
chunk reads only values that are final before its level."""
import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT / "geos-chem_amd" / "mechgen"))


def raw_mech(m):
    return {"nvar": m.nvar, "nfix": m.nfix, "nspec": m.nspec, "nreact": m.nreact,
            "lu_nonzero": m.lu_nonzero, "reactions": m.reactions, "lu_irow": list(m.lu_irow),
            "lu_icol": list(m.lu_icol), "lu_crow": list(m.lu_crow), "lu_diag": list(m.lu_diag)}


def programs(m):
    from emit_tables import build_tables
    import wave_sched as W
    raw = raw_mech(m)
    t = build_tables(raw)
    return W, t, W.build_wave_program(raw, t)


def fresh_lds(prog, rng):
    L = prog["layout"]
    sh = list(rng.uniform(-1.0, 1.0, L["SH_SIZE"]))
    sh[L["ZERO"]] = 0.0
    for k, c in enumerate(prog["coef"]):
        sh[L["COEF0"] + k] = c
    return sh, L


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
def test_levels_read_only_final_values(name, mech_of):
    W, t, prog = programs(mech_of(name))
    for P in ("fv", "jv", "lu", "sf", "sb"):
        assert W.level_dependencies_ok(prog, P), P


def test_dependency_check_catches_merged_levels(fullchem):
    """Merging two consecutive levels (dropping the ordering point between
    them) makes chunks read values written in their own level: the check
    fails for every phase with more than one level."""
    import copy
    W, t, prog = programs(fullchem)
    for P in ("lu", "sf", "sb"):
        bad = copy.deepcopy(prog)
        (lev, rs, head, term), pad = bad[P]
        del lev[2]
        assert not W.level_dependencies_ok(bad, P), P


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
def test_programs_reproduce_table_walk(name, mech_of):
    m = mech_of(name)
    W, t, prog = programs(m)
    rng = np.random.default_rng(3)
    nv, nz = m.nvar, m.lu_nonzero
    crow, diag, icol = m.lu_crow, m.lu_diag, m.lu_icol

    # Vdot: A in the scratch, rows into X
    sh, L = fresh_lds(prog, rng)
    AB0, X0 = L["AB0"], L["X0"]
    ref = []
    for i in range(nv):
        acc = 0.0
        for k in range(t["fv_ptr"][i], t["fv_ptr"][i + 1]):
            acc = acc + float(t["fv_c"][k]) * sh[AB0 + t["fv_r"][k]]
        ref.append(acc)

    def st_plain(s, d, hi, acc):
        s[d] = acc
    W.run_phase(prog, "fv", sh, lambda acc, a, b: acc + b * a, False, st_plain)
    assert sh[X0:X0 + nv] == ref

    # Jacobian: partials in the scratch, Ghimj = -JVS + diag into the LU block
    sh, L = fresh_lds(prog, rng)
    dg = 37.5
    ref = []
    dset = set(diag[:nv])
    for q in range(nz):
        acc = 0.0
        for k in range(t["jv_ptr"][q], t["jv_ptr"][q + 1]):
            acc = acc + float(t["jv_c"][k]) * sh[AB0 + t["jv_b"][k]]
        v = -1.0 * acc
        if q in dset:
            v = v + dg
        ref.append(v)

    def st_jac(s, d, hi, acc):
        v = -1.0 * acc
        if hi:
            v = v + dg
        s[d] = v
    W.run_phase(prog, "jv", sh, lambda acc, a, b: acc + b * a, False, st_jac)
    assert sh[:nz] == ref

    # LU (row-Crout, the lane solvers' cr tables) on a diagonally dominant matrix
    sh, L = fresh_lds(prog, rng)
    for k in range(nv):
        sh[diag[k]] += 40.0
    lu = list(sh[:nz])
    for q in range(nz):
        acc = lu[q]
        for k in range(t["cr_ptr"][q], t["cr_ptr"][q + 1]):
            acc = acc + (-lu[t["cr_l"][k]]) * lu[t["cr_u"][k]]
        if t["cr_div"][q] >= 0:
            acc = acc / lu[t["cr_div"][q]]
        lu[q] = acc

    def st_div(s, d, hi, acc):
        s[d] = acc / s[hi - 1] if hi else acc
    W.run_phase(prog, "lu", sh, lambda acc, a, b: acc + (-a) * b, True, st_div)
    assert sh[:nz] == lu

    # substitutions with that LU (KppSolve order)
    x = list(sh[X0:X0 + nv])
    for i in range(nv):
        acc = x[i]
        for k in range(crow[i], diag[i]):
            acc = acc + (-lu[k]) * x[icol[k]]
        x[i] = acc
    for i in range(nv - 1, -1, -1):
        acc = x[i]
        for k in range(diag[i] + 1, crow[i + 1]):
            acc = acc + (-lu[k]) * x[icol[k]]
        x[i] = acc / lu[diag[i]]
    W.run_phase(prog, "sf", sh, lambda acc, a, b: acc + (-a) * b, True, st_plain)
    W.run_phase(prog, "sb", sh, lambda acc, a, b: acc + (-a) * b, True, st_div)
    assert sh[X0:X0 + nv] == x


def test_emitted_tables_match_programs(fullchem):
    """The generated header holds the programs the generator builds now."""
    W, t, prog = programs(fullchem)
    hdr = (ROOT / "geos-chem_amd" / "csrc" / "gen" / "mech_fullchem.h").read_text()
    L = prog["layout"]
    assert "WV_X0 = %d," % L["X0"] in hdr and "WV_SH = %d;" % L["SH_SIZE"] in hdr
    for name, ctype, vals in W.emit_arrays(prog):
        assert "%s %s[%d]" % (ctype, name, len(vals)) in hdr, name


@pytest.mark.parametrize("name", ["fullchem", "Hg", "carbon"])
def test_program_indices_inside_lds_block(name, mech_of):
    """Every slot a program names (destination, divisor, both operands of
    every step, pads included) lies inside the mechanism's LDS block, and
    destinations never touch the constant slots (ZERO, coefficients)."""
    W, t, prog = programs(mech_of(name))
    L = prog["layout"]
    for P in ("fv", "jv", "lu", "sf", "sb"):
        if P not in prog:
            continue
        (lev, rs, head, term_rows), pad = prog[P]
        assert len(head) == 64 * (len(rs) - 1) and len(term_rows) == rs[-1]
        assert lev[0] == 0 and lev[-1] == len(rs) - 1 and list(lev) == sorted(lev)
        for h in head:
            d, hi = h & 0xFFFF, h >> 16
            if d == W.IDLE:
                continue
            assert d < L["ZERO"] or L["X0"] <= d < L["COEF0"]
            if P in ("lu", "sb") and hi:
                assert hi - 1 < L["ZERO"]
        for row in W._encode_terms(term_rows, pad):
            assert (row & 0xFFFF) < L["SH_SIZE"] and (row >> 16) < L["SH_SIZE"]
