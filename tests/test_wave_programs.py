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
    """Random LDS block (the +0.0 slot set) and reaction scratch (its +0.0
    pad slot set) and the coefficient table."""
    L = prog["layout"]
    sh = list(rng.uniform(-1.0, 1.0, L["SH_SIZE"]))
    sh[L["ZERO"]] = 0.0
    sh[L["ONE"]] = 1.0
    ab = list(rng.uniform(-1.0, 1.0, L["NAB"] + 1))
    ab[L["NAB"]] = 0.0
    return sh, L, ab, list(prog["coef"])


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
    sh, L, ab, coef = fresh_lds(prog, rng)
    X0 = L["X0"]
    ref = []
    for i in range(nv):
        acc = 0.0
        for k in range(t["fv_ptr"][i], t["fv_ptr"][i + 1]):
            acc = acc + float(t["fv_c"][k]) * ab[t["fv_r"][k]]
        ref.append(acc)

    def st_plain(s, d, hi, acc):
        s[d] = acc
    W.run_phase(prog, "fv", sh, lambda acc, a, b: acc + b * a, False, st_plain, ab=ab, coef=coef)
    assert sh[X0:X0 + nv] == ref

    # Jacobian: partials in the scratch, Ghimj = -JVS + diag into the LU block
    sh, L, ab, coef = fresh_lds(prog, rng)
    dg = 37.5
    ref = []
    dset = set(diag[:nv])
    for q in range(nz):
        acc = 0.0
        for k in range(t["jv_ptr"][q], t["jv_ptr"][q + 1]):
            acc = acc + float(t["jv_c"][k]) * ab[t["jv_b"][k]]
        v = -1.0 * acc
        if q in dset:
            v = v + dg
        ref.append(v)

    def st_jac(s, d, hi, acc):
        v = -1.0 * acc
        if hi:
            v = v + dg
        s[d] = v
    W.run_phase(prog, "jv", sh, lambda acc, a, b: acc + b * a, False, st_jac, ab=ab, coef=coef)
    assert sh[:nz] == ref

    # LU (row-Crout, the lane solvers' cr tables) on a diagonally dominant matrix
    sh, L, _, _ = fresh_lds(prog, rng)
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
    arrays, consts = W.emit_arrays(prog)
    for name, ctype, vals in arrays:
        assert "%s %s[%d]" % (ctype, name, len(vals)) in hdr, name
    for name, v in consts:
        assert "%s = %d" % (name, v) in hdr, name


@pytest.mark.parametrize("name", ["fullchem", "Hg", "carbon"])
def test_program_offsets_inside_their_blocks(name, mech_of):
    """Every byte offset an encoded program names -- header destination and
    divisor, both operands of every op, pads included -- lies inside its
    block: the LDS block for LU and substitution operands and every
    destination, the reaction scratch (with its pad slot) and the
    coefficient table for the Vdot and Jacobian operands; destinations
    never touch the +0.0 or +1.0 slots or the state Y, and an idle lane's
    destination is its own sink slot; the block flags mark every round's
    first and last block; the ops array holds PF + XPAD pad blocks past NB,
    NB a multiple of PF, and the flag words cover them."""
    W, t, prog = programs(mech_of(name))
    L = prog["layout"]
    sh8, ab8, cf8 = 8 * L["SH_SIZE"], 8 * (L["NAB"] + 1), 8 * len(prog["coef"])
    sink = 8 * L["SINK"]
    for P in ("fv", "jv", "lu", "sf", "sb"):
        if P not in prog:
            continue
        ops, flag, nb, nr = W.encode(prog, P)
        nblk = nb + W.PF + W.XPAD
        assert nb % W.PF == 0 and len(ops) == nblk * 2 * 64 * 4
        assert len(flag) * 8 >= nblk
        fl = [(flag[k // 8] >> (4 * (k % 8))) & 15 for k in range(len(flag) * 8)]
        real = max(k for k in range(nblk) if fl[k] & 2) + 1     # blocks of the rounds
        assert sum(f & 1 for f in fl[:real]) == nr and sum((f >> 1) & 1 for f in fl) == nr
        # the padding to NB starts a round of sink headers (no end); past NB nothing
        assert fl[real:nb] == ([1] + [0] * (nb - real - 1) if nb > real else [])
        assert all(f == 0 for f in fl[nb:])
        lds = P not in ("fv", "jv")
        start = False
        for k in range(nblk):
            f = fl[k]
            for ln in range(64):
                w = [ops[((2 * k + h) * 64 + ln) * 4 + j] for h in range(2) for j in range(4)]
                u = [(w[0], w[1]), (w[2], w[3]), (w[4], w[5]), (w[6], w[7])]
                if f & 1:
                    d, x = u[0]
                    assert d % 8 == 0
                    if k >= real:
                        assert d == sink + 8 * ln
                    assert (d < 8 * L["ZERO"] or 8 * L["X0"] <= d < 8 * L["Y0"] or d == sink + 8 * ln)
                    if lds:
                        assert x % 8 == 0 and (x < 8 * L["ZERO"] or x == 8 * L["ONE"])
                    else:
                        assert x in (0, 1)
                    u = u[1:]
                for a, b in u:
                    assert a % 8 == 0 and b % 8 == 0
                    if lds:
                        assert a < sh8 and b < sh8 and a < sink and b < sink
                    else:
                        assert a < ab8 and b < cf8
