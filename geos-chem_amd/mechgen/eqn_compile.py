"""Mechanism compiler from KPP source: reads a mechanism's .kpp file (and the
.eqn file it includes) and writes this project's description of the
mechanism (the JSON of kpp_parse.py: species, reactions with reactants, net
stoichiometry and rate expressions, the sparse LU pattern), without the KPP
generator.

Input language (the subset GEOS-Chem's mechanisms use; KPP/fullchem/
fullchem.kpp, KPP/fullchem/fullchem.eqn:121-2031, KPP/Hg/Hg.kpp,
KPP/carbon/carbon.kpp and their .eqn files):
  { ... }                comments (may span lines); // to the end of a line
  #INCLUDE file          another source file (relative to the current one)
  #DEFVAR / #DEFFIX      "NAME = <atoms>;" variable / fixed species
  #EQUATIONS             "lhs = rhs : rate;" -- terms "[c]NAME" joined by
                         + or - (a product with a minus sign is consumed);
                         "hv" is a placeholder; the rate is a Fortran
                         expression, copied as it is
  #FAMILIES              "NAME : [c]MEMBER + ...;" prod/loss families: a
                         family whose name starts with P gets, in every
                         reaction, the net production of its members (the
                         member coefficients times the members' net
                         stoichiometry) as a product when that is positive;
                         one whose name starts with L gets the net
                         consumption likewise.  The family species are new
                         variable species, appended to #DEFVAR's.  A declared
                         species that no equation uses is dropped.
  #INTEGRATOR            feuler: no Jacobian (no LU pattern)
  #INLINE ... #ENDINLINE, other # directives: skipped.

Species order.  The C ABI keeps KPP's order (the ind_* of
gckpp_Parameters.F90), so for GEOS-Chem's own mechanisms the order is read
from that file when it is given; a mechanism without KPP output (a user
mechanism) gets this compiler's own order: Markowitz pivoting on the
Jacobian's structure (at each step the remaining variable species whose
row count times column count, less the pivot itself, is smallest; ties to
the first in #DEFVAR order), which keeps the LU fill small.  Fixed species
follow the variable ones in #DEFFIX order.

Sparse pattern: the Jacobian J(i,j) is structurally nonzero where species j
is a reactant of a reaction that changes species i, plus the diagonal; the
LU pattern is its symbolic row-wise elimination (no pivoting), stored
row-wise with a diagonal pointer, the layout of KPP's LU_IROW / LU_ICOL /
LU_CROW / LU_DIAG (KPP/fullchem/gckpp_JacobianSP.F90).

Usage:  python eqn_compile.py <mech.kpp> <out.json> [--order gckpp_Parameters.F90] [--name NAME]
"""
import json
import re
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE))
from kpp_parse import fortran_expr_to_c  # noqa: E402


# ---------------------------------------------------------------------------
# lexing
def _strip_comments(text):
    """Remove { } comments (possibly multi-line) and // line comments."""
    out, depth, i = [], 0, 0
    while i < len(text):
        c = text[i]
        if depth == 0 and text.startswith("//", i):
            j = text.find("\n", i)
            i = len(text) if j < 0 else j
            continue
        if c == "{":
            depth += 1
        elif c == "}" and depth > 0:
            depth -= 1
        elif depth == 0:
            out.append(c)
        i += 1
    return "".join(out)


def _read_source(path, seen=None):
    """The text of a KPP source file with its #INCLUDEs expanded (#include
    of the atom table, atoms.kpp, is dropped: atoms only check balance)."""
    path = Path(path)
    seen = seen or set()
    if path.resolve() in seen:
        raise ValueError("recursive #INCLUDE of %s" % path)
    seen.add(path.resolve())
    text = _strip_comments(path.read_text())
    text = re.sub(r"#INLINE.*?#ENDINLINE", "", text, flags=re.S | re.I)
    out = []
    for line in text.splitlines():
        m = re.match(r"\s*#INCLUDE\s+(\S+)", line, re.I)
        if m:
            name = m.group(1)
            if name.lower() == "atoms.kpp" or name.lower().endswith("/atoms.kpp"):
                continue
            inc = path.parent / name
            if not inc.exists():
                raise FileNotFoundError("#INCLUDE %s (from %s)" % (name, path))
            out.append(_read_source(inc, seen))
        else:
            out.append(line)
    return "\n".join(out)


def _sections(text):
    """Split the expanded source into (directive, body) pairs; #INLINE blocks
    are dropped."""
    text = re.sub(r"#INLINE.*?#ENDINLINE", "", text, flags=re.S | re.I)
    parts = re.split(r"(?m)^\s*(#[A-Za-z_0-9]+)", text)
    out = []
    for k in range(1, len(parts), 2):
        out.append((parts[k].upper(), parts[k + 1]))
    return out


_TERM = re.compile(r"^([0-9]*\.?[0-9]*(?:[eEdD][+-]?[0-9]+)?)\s*([A-Za-z_][A-Za-z_0-9]*)$")


def _terms(side):
    """[(name, coefficient)] of one side of an equation ('-' terms negative)."""
    side = side.strip()
    if not side:
        return []
    out = []
    # split on + and - that separate terms (not exponent signs)
    toks = re.split(r"(?<![eEdD0-9.][eEdD])\s*([+-])\s*", side)
    sign = 1.0
    if toks and toks[0] == "":
        toks = toks[1:]
    for t in toks:
        t = t.strip()
        if t in ("+", "-"):
            sign = -1.0 if t == "-" else 1.0
            continue
        if not t:
            continue
        m = _TERM.match(t.replace(" ", ""))
        if not m:
            raise ValueError("cannot read the term %r in %r" % (t, side))
        c = m.group(1)
        coef = float(c.replace("d", "e").replace("D", "e")) if c not in ("", ".") else 1.0
        out.append((m.group(2), sign * coef))
        sign = 1.0
    return out


def parse(kpp_path):
    """Species, families and equations of a KPP mechanism source."""
    text = _read_source(kpp_path)
    defvar, deffix, eqns, fams = [], [], [], []
    for d, body in _sections(text):
        # a stray separator between two comments after an entry (",") is
        # not part of the next entry
        entries = [e.strip().lstrip(",").strip() for e in body.split(";")]
        if d == "#DEFVAR" or d == "#DEFFIX":
            for e in entries:
                if not e:
                    continue
                name = e.split("=")[0].strip()
                (defvar if d == "#DEFVAR" else deffix).append(name)
        elif d == "#EQUATIONS":
            for e in entries:
                if not e:
                    continue
                e = re.sub(r"^<[^>]*>", "", e).strip()      # an equation tag
                lhs_rhs, rate = e.split(":", 1)
                lhs, rhs = lhs_rhs.split("=", 1)
                eqns.append((_terms(lhs), _terms(rhs), " ".join(rate.split())))
        elif d == "#FAMILIES":
            for e in entries:
                if not e:
                    continue
                name, members = e.split(":", 1)
                fams.append((name.strip(), _terms(members)))
    m = re.search(r"(?m)^\s*#INTEGRATOR\s+(\S+)", text, re.I)
    return dict(defvar=defvar, deffix=deffix, equations=eqns, families=fams,
                integrator=m.group(1).lower() if m else "rosenbrock")


# ---------------------------------------------------------------------------
# reactions
def _printed(v):
    """A stoichiometric coefficient as KPP writes it into the generated code:
    six significant digits (1 - 0.79 is 0.21, not 0.20999999999999996)."""
    return float("%g" % v)


def reactions_by_name(p):
    """[(reactants {name: count}, net stoichiometry {name: coef}, rate)] in
    equation order, with the family species added."""
    out = []
    for lhs, rhs, rate in p["equations"]:
        react, net = {}, {}
        for n, c in lhs:
            if n.lower() == "hv":
                continue
            react[n] = react.get(n, 0) + int(round(c))
            net[n] = net.get(n, 0.0) - c
        for n, c in rhs:
            if n.lower() == "hv":
                continue
            net[n] = net.get(n, 0.0) + c
        fam_add = {}
        for fname, members in p["families"]:
            # products first, then the reactants; a total within 1e-9 of
            # zero is zero (0.9 + 0.1 - 1)
            mem = dict(members)
            prod = sum(mem[n] * c for n, c in rhs if n in mem)
            cons = sum(mem[n] * c for n, c in lhs if n in mem)
            tot = _printed(prod - cons)
            if abs(tot) < 1e-9:
                continue
            if fname.upper().startswith("P") and tot > 0:
                fam_add[fname] = tot
            elif fname.upper().startswith("L") and tot < 0:
                fam_add[fname] = -tot
        for n, c in fam_add.items():
            net[n] = net.get(n, 0.0) + c
        net = {n: _printed(c) for n, c in net.items()}
        net = {n: c for n, c in net.items() if c != 0.0}
        out.append((react, net, rate))
    return out


def species_lists(p, rx):
    """(variable species in #DEFVAR order plus the families, fixed species);
    a declared species that no equation uses is dropped, as KPP does."""
    used = set()
    for react, net, _ in rx:
        used.update(react)
        used.update(net)
    var = [s for s in p["defvar"] if s in used]
    for fname, _ in p["families"]:
        if fname not in var:
            var.append(fname)
    fix = [s for s in p["deffix"] if s not in var and s in used]
    return var, fix


def jacobian_structure(rx, index, nvar):
    """Sets of columns per row: J(i, j) nonzero where variable species j is a
    reactant of a reaction that changes variable species i; the diagonal."""
    rows = [set([i]) for i in range(nvar)]
    for react, net, _ in rx:
        cols = [index[n] for n in react if index[n] < nvar]
        for n, c in net.items():
            i = index[n]
            if i < nvar:
                rows[i].update(cols)
    return rows


def markowitz_order(rows_by_name_index, nvar):
    """Pivot order by the Markowitz criterion on the structure (symmetric
    permutation, diagonal pivots): repeatedly the remaining species whose
    (row count - 1) * (column count - 1) in the remaining submatrix, with
    the fill of the previous pivots, is smallest (the first in the input
    order on ties)."""
    rows = [set(r) for r in rows_by_name_index]
    cols = [set() for _ in range(nvar)]
    for i, r in enumerate(rows):
        for j in r:
            cols[j].add(i)
    left = set(range(nvar))
    order = []
    while left:
        best, bk = None, None
        for k in sorted(left):
            key = (len(rows[k] & left) - 1) * (len(cols[k] & left) - 1)
            if bk is None or key < bk:
                best, bk = k, key
        k = best
        left.discard(k)
        order.append(k)
        rk = rows[k] & left
        ck = cols[k] & left
        for i in ck:          # eliminating k fills row i with row k's columns
            for j in rk:
                if j not in rows[i]:
                    rows[i].add(j)
                    cols[j].add(i)
    return order


def symbolic_lu(rows, nvar):
    """Row-wise symbolic elimination without pivoting: the LU pattern as
    (lu_irow, lu_icol, lu_crow, lu_diag), 0-based, rows in order, columns
    ascending in each row."""
    pat = [set(r) for r in rows]
    for i in range(nvar):
        # row i is eliminated with every earlier row k it has a nonzero in
        done = set()
        while True:
            ks = sorted(k for k in pat[i] if k < i and k not in done)
            if not ks:
                break
            k = ks[0]
            done.add(k)
            pat[i].update(c for c in pat[k] if c > k)
    irow, icol, crow, diag = [], [], [], []
    for i in range(nvar):
        crow.append(len(icol))
        for c in sorted(pat[i]):
            if c == i:
                diag.append(len(icol))
            irow.append(i)
            icol.append(c)
    crow.append(len(icol))
    diag.append(len(icol))
    return irow, icol, crow, diag


def kpp_order(parameters_path):
    """Species names in the order of KPP's ind_* (gckpp_Parameters.F90)."""
    txt = Path(parameters_path).read_text()
    idx = {m.group(1): int(m.group(2)) for m in
           re.finditer(r"INTEGER, PARAMETER :: ind_(\w+) = (\d+)", txt)}
    names = [None] * max(idx.values())
    for n, i in idx.items():
        names[i - 1] = n
    return names


def _is_literal(rate):
    """A rate written as a plain number (KPP's own number syntax, no Fortran
    'd' exponent or kind suffix): KPP writes it into Fun as a constant and
    leaves RCONST unset for the reaction."""
    return re.fullmatch(r"[0-9]*\.?[0-9]+(?:[eE][+-]?[0-9]+)?", rate.strip()) is not None


def compile_mechanism(kpp_path, name=None, order_from=None):
    """The mechanism description (kpp_parse.py's JSON schema) of a KPP
    source.  order_from: a gckpp_Parameters.F90 whose species order is kept
    (KPP's own, for the C ABI); otherwise the Markowitz order."""
    p = parse(kpp_path)
    rx = reactions_by_name(p)
    var, fix = species_lists(p, rx)
    nvar = len(var)
    if order_from is not None:
        names = kpp_order(order_from)
        if sorted(names) != sorted(var + fix):
            missing = set(var + fix) ^ set(names)
            raise ValueError("species of %s and %s differ: %s" % (kpp_path, order_from, sorted(missing)))
        if set(names[:nvar]) != set(var):
            raise ValueError("variable species of %s are not KPP's first %d" % (order_from, nvar))
        species = names
    else:
        index0 = {n: i for i, n in enumerate(var + fix)}
        rows0 = jacobian_structure(rx, index0, nvar)
        order = markowitz_order(rows0, nvar)
        species = [var[k] for k in order] + fix
    index = {n: i for i, n in enumerate(species)}
    if p["integrator"] == "feuler":
        # the forward-Euler integrator needs no Jacobian (KPP writes a
        # one-entry placeholder size and no pattern)
        nonzero, irow, icol, crow, diag = 1, [], [], [], []
    else:
        rows = jacobian_structure(rx, index, nvar)
        nonzero = sum(len(r) for r in rows)
        irow, icol, crow, diag = symbolic_lu(rows, nvar)
    reactions, eqn_names = [], []
    for react, net, rate in rx:
        lit = _is_literal(rate)
        if not any(index[n] < nvar for n in net):
            # changes no variable species: nothing in Fun or the Jacobian
            # (KPP drops the reaction's product term)
            react = {}
        reactions.append({
            "reactants": sorted([[index[n], c] for n, c in react.items()]),
            # the ODE covers the variable species only
            "stoich": sorted([[index[n], c] for n, c in net.items() if index[n] < nvar]),
            "rate": rate.strip().lower() if lit else fortran_expr_to_c("(%s)" % rate.replace(" ", "")),
            "inlined": lit})
        lhs = " + ".join(("%d%s" % (c, n) if c != 1 else n) for n, c in react.items())
        rhs = " + ".join(("%g%s" % (c, n) if c != 1 else n) for n, c in net.items() if c > 0)
        eqn_names.append("%s --> %s" % (lhs, rhs))
    return {"name": name or Path(kpp_path).stem, "source": "%s (eqn_compile.py)" % Path(kpp_path).name,
            "nspec": len(species), "nvar": nvar, "nfix": len(species) - nvar,
            "nreact": len(reactions), "nonzero": nonzero, "lu_nonzero": max(len(icol), 1),
            "species": species, "reactions": reactions, "lu_irow": irow, "lu_icol": icol,
            "lu_crow": crow, "lu_diag": diag, "eqn_names": eqn_names}


def main(argv):
    import argparse
    ap = argparse.ArgumentParser()
    ap.add_argument("kpp")
    ap.add_argument("out")
    ap.add_argument("--order", default=None, help="gckpp_Parameters.F90 whose species order to keep")
    ap.add_argument("--name", default=None)
    a = ap.parse_args(argv)
    d = compile_mechanism(a.kpp, a.name, a.order)
    Path(a.out).write_text(json.dumps(d, separators=(",", ":")))
    print("%s: NSPEC=%d NVAR=%d NREACT=%d NONZERO=%d LU_NONZERO=%d" %
          (d["name"], d["nspec"], d["nvar"], d["nreact"], d["nonzero"], d["lu_nonzero"]))


if __name__ == "__main__":
    main(sys.argv[1:])
