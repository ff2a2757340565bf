"""Mechanism compiler, front end: reads a KPP-3.2.0-generated mechanism and
writes this project's own compact description of it (JSON).

It extracts only the mechanism *tables* -- species order, fixed species,
reactant lists, net stoichiometry, rate-coefficient expressions and the sparse
LU pattern -- which every solver for the mechanism must share so that it is a
drop-in for the reference call.  No solver code is taken over: the kernels are
generated from these tables by ``emit_tables.py`` and ``emit_rates.py``, and the
oracle reads the JSON.

Sources read (paths relative to the reference checkout, KPP/<mech>/):
  gckpp_Parameters.F90   NSPEC/NVAR/NFIX/NREACT/LU_NONZERO, ind_<name>
                         (e.g. KPP/fullchem/gckpp_Parameters.F90:33-58,62-434)
  gckpp_Function.F90     SUBROUTINE Fun: "A(r) = RCT(r)*V(i)*F(j)" gives the
                         reactants of reaction r (or a literal that KPP inlined
                         in place of RCT(r)); "Vdot(i) = c*A(r) ..." gives the
                         net stoichiometry (KPP/fullchem/gckpp_Function.F90:51-2152)
  gckpp_Rates.F90        SUBROUTINE Update_RCONST: "RCONST(r) = (expr)"
                         (KPP/fullchem/gckpp_Rates.F90:408-1503)
  gckpp_JacobianSP.F90   LU_IROW, LU_ICOL, LU_CROW, LU_DIAG
                         (KPP/fullchem/gckpp_JacobianSP.F90:35-1093)
  gckpp_Monitor.F90      EQN_NAMES, for labels only (KPP/fullchem/gckpp_Monitor.F90:158-1252)

Usage:  python kpp_parse.py <reference-root> <mech> <out.json>
"""
import json
import re
import sys
from pathlib import Path


def _read_joined(path):
    """Return the file's lines with Fortran '&' continuations joined and
    comments removed."""
    out, cur = [], ""
    for raw in Path(path).read_text().splitlines():
        line = raw.rstrip()
        s = line.strip()
        if not cur and (s.startswith("!") or not s):
            continue
        # strip trailing comment (no '!' inside strings in these files' code lines)
        if "!" in line and "'" not in line and '"' not in line:
            line = line[: line.index("!")].rstrip()
            s = line.strip()
        if cur:
            if s.startswith("&"):
                s = s[1:]
            cur += s
        else:
            cur = line.strip()
        if cur.endswith("&"):
            cur = cur[:-1]
            continue
        out.append(cur)
        cur = ""
    if cur:
        out.append(cur)
    return out


def _fortran_arrays(text, name):
    """Collect integer array constants NAME_<k> = (/ ... /) (or NAME = (/.../))
    in order, concatenated."""
    parts = {}
    pat = re.compile(r"INTEGER, PARAMETER, DIMENSION\(\d+\) :: (%s(?:_\d+)?) = \(/(.*?)/\)" % name,
                     re.S)
    for m in pat.finditer(text):
        nm, body = m.group(1), m.group(2)
        body = re.sub(r"!.*", "", body)
        vals = [int(v) for v in re.findall(r"-?\d+", body.replace("&", " "))]
        parts[nm] = vals
    if name in parts and not any(k.startswith(name + "_") for k in parts):
        return parts[name]
    keys = sorted((k for k in parts if k.startswith(name + "_")), key=lambda k: int(k.split("_")[-1]))
    res = []
    for k in keys:
        res += parts[k]
    return res


def parse_parameters(path):
    txt = Path(path).read_text()
    sizes = {}
    for k in ("NSPEC", "NVAR", "NFIX", "NREACT", "NONZERO", "LU_NONZERO", "NFAM"):
        m = re.search(r"INTEGER, PARAMETER :: %s = (\d+)" % k, txt)
        sizes[k] = int(m.group(1)) if m else 0
    ind = {}
    for m in re.finditer(r"INTEGER, PARAMETER :: ind_(\w+) = (\d+)", txt):
        ind[m.group(1)] = int(m.group(2))
    names = [None] * sizes["NSPEC"]
    for nm, i in ind.items():
        names[i - 1] = nm
    assert all(n is not None for n in names), "species index gap"
    return sizes, names


_A_RE = re.compile(r"^A\((\d+)\) = (.*)$")
_VDOT_RE = re.compile(r"^Vdot\((\d+)\) = (.*)$")


def parse_function(path, nvar, nreact):
    """Reactants and net stoichiometry from the aggregate Fun routine."""
    lines = _read_joined(path)
    # restrict to SUBROUTINE Fun ( ... ) .. END SUBROUTINE Fun
    start = next(i for i, l in enumerate(lines) if re.match(r"SUBROUTINE Fun \(", l))
    end = next(i for i in range(start, len(lines)) if lines[i].startswith("END SUBROUTINE Fun"))
    body = lines[start:end]
    reactants = [None] * nreact
    const = [None] * nreact
    stoich = [[] for _ in range(nreact)]
    for l in body:
        l = re.sub(r"^IF \(DO_FUN\(\d+\)\)", "", l).strip()
        m = _A_RE.match(l)
        if m:
            r = int(m.group(1)) - 1
            facs = m.group(2).split("*")
            head = facs[0]
            if head.startswith("RCT("):
                assert head == "RCT(%d)" % (r + 1), l
            else:
                const[r] = head  # literal rate inlined by KPP
            cnt = {}
            for f in facs[1:]:
                mv = re.match(r"^([VF])\((\d+)\)$", f)
                assert mv, l
                idx = int(mv.group(2)) - 1 + (nvar if mv.group(1) == "F" else 0)
                cnt[idx] = cnt.get(idx, 0) + 1
            reactants[r] = sorted(cnt.items())
            continue
        m = _VDOT_RE.match(l)
        if m:
            i = int(m.group(1)) - 1
            expr = m.group(2).replace(" ", "")
            for sign, coef, rr in re.findall(r"([+-]?)(?:([0-9.eEdD+-]+?)\*)?A\((\d+)\)", expr):
                c = float(coef.replace("d", "e").replace("D", "e")) if coef else 1.0
                if sign == "-":
                    c = -c
                stoich[int(rr) - 1].append([i, c])
    for r in range(nreact):
        if reactants[r] is None:
            # KPP emits no A(r) for a reaction that changes no variable
            # species (e.g. a photolysis among fixed species): inert here.
            assert not stoich[r], "reaction %d has stoichiometry but no rate" % (r + 1)
            reactants[r] = []
    return reactants, const, stoich


_FUNC_CALL = re.compile(r"[A-Za-z_][A-Za-z_0-9]*")


def fortran_expr_to_c(expr):
    """Translate one rate expression from Fortran to C.

    Numeric literals become C doubles; PHOTOL(n), K_MT(n), K_CLD(n) become
    0-based array reads of the per-cell state; C(ind_X) and SR_MW(ind_X) keep
    the species name (resolved by the emitter); State_Het stays a pointer.
    """
    s = expr.strip()
    # numbers: 1.0d-12, 2.d0, 1.0E8_dp, 3_dp, 4.51e-12
    def num(m):
        t = m.group(0)
        t = re.sub(r"_dp$", "", t)
        t = t.replace("d", "e").replace("D", "e")
        if "e" not in t and "." not in t:
            t += ".0"
        return t
    s = re.sub(r"PHOTOL\((\d+)\)", lambda m: "PHOTOL[%d]" % (int(m.group(1)) - 1), s)
    s = re.sub(r"K_MT\((\d+)\)", lambda m: "K_MT[%d]" % (int(m.group(1)) - 1), s)
    s = re.sub(r"K_CLD\((\d+)\)", lambda m: "K_CLD[%d]" % (int(m.group(1)) - 1), s)
    s = re.sub(r"\bC\(ind_(\w+)\)", r"CONC(\1)", s)
    s = re.sub(r"\bSR_MW\(ind_(\w+)\)", r"SRMW(\1)", s)
    s = re.sub(r"\bexp\(", "exp(", s, flags=re.I)
    s = re.sub(r"(?<![A-Za-z_0-9.\[])(\d+\.?\d*|\.\d+)([dDeE][+-]?\d+)?(_dp)?(?![A-Za-z_0-9(\]])", num, s)
    s = s.replace(".TRUE.", "1").replace(".FALSE.", "0")
    assert "**" not in s, s
    return s


def parse_rates(path, nreact):
    lines = _read_joined(path)
    start = next(i for i, l in enumerate(lines) if l.startswith("SUBROUTINE Update_RCONST"))
    end = next(i for i in range(start, len(lines)) if lines[i].upper().startswith("END SUBROUTINE UPDATE_RCONST"))
    rates = [None] * nreact
    raw = [None] * nreact
    for l in lines[start:end]:
        m = re.match(r"^RCONST\((\d+)\) = (.*)$", l)
        if m:
            r = int(m.group(1)) - 1
            raw[r] = m.group(2)
            rates[r] = fortran_expr_to_c(m.group(2))
    return rates, raw


def parse_jacsp(path):
    txt = Path(path).read_text()
    return {k: _fortran_arrays(txt, k) for k in ("LU_IROW", "LU_ICOL", "LU_CROW", "LU_DIAG")}


def parse_eqn_names(path, nreact):
    txt = Path(path).read_text()
    names = re.findall(r"'([^']*)'", txt[txt.find("EQN_NAMES"):])
    out = []
    for n in names:
        n = n.strip()
        if "-->" in n:
            out.append(n)
    return out[:nreact] if len(out) >= nreact else None


def build(refroot, mech):
    d = Path(refroot) / "KPP" / mech
    sizes, names = parse_parameters(d / "gckpp_Parameters.F90")
    nvar, nspec, nreact = sizes["NVAR"], sizes["NSPEC"], sizes["NREACT"]
    reactants, const, stoich = parse_function(d / "gckpp_Function.F90", nvar, nreact)
    rates, raw = parse_rates(d / "gckpp_Rates.F90", nreact)
    if (d / "gckpp_JacobianSP.F90").exists():
        lu = parse_jacsp(d / "gckpp_JacobianSP.F90")
    else:  # e.g. carbon: forward Euler, KPP generated no sparse Jacobian
        lu = {"LU_IROW": [], "LU_ICOL": [], "LU_CROW": [], "LU_DIAG": []}
    for k in lu:
        lu[k] = [v - 1 for v in lu[k]]
    reactions = []
    for r in range(nreact):
        c = const[r]
        rate = rates[r]
        if c is not None:
            # KPP replaced RCT(r) by this literal everywhere; RCONST(r) is unset.
            rate = fortran_expr_to_c(c)
        assert rate is not None, "reaction %d has no rate" % (r + 1)
        reactions.append({
            "reactants": [[i, m] for i, m in reactants[r]],
            "stoich": stoich[r],
            "rate": rate,
            "inlined": c is not None,
        })
    eqn = parse_eqn_names(d / "gckpp_Monitor.F90", nreact)
    mechd = {
        "name": mech,
        "source": "KPP/%s (KPP-3.2.0 generated tables)" % mech,
        "nspec": nspec, "nvar": nvar, "nfix": sizes["NFIX"], "nreact": nreact,
        "nonzero": sizes["NONZERO"], "lu_nonzero": sizes["LU_NONZERO"],
        "species": names,
        "reactions": reactions,
        "lu_irow": lu["LU_IROW"], "lu_icol": lu["LU_ICOL"],
        "lu_crow": lu["LU_CROW"], "lu_diag": lu["LU_DIAG"],
        "eqn_names": eqn,
    }
    if mechd["lu_irow"]:
        assert len(mechd["lu_irow"]) == mechd["lu_nonzero"]
        assert len(mechd["lu_crow"]) == nvar + 1
    return mechd


if __name__ == "__main__":
    root, mech, out = sys.argv[1], sys.argv[2], sys.argv[3]
    m = build(root, mech)
    Path(out).write_text(json.dumps(m, separators=(",", ":")))
    print("%s: nvar=%d nfix=%d nreact=%d lu_nonzero=%d" % (mech, m["nvar"], m["nfix"], m["nreact"], m["lu_nonzero"]))
