"""Mechanism compiler, back end: writes the device tables one mechanism needs
(csrc/gen/mech_<name>.h) from mech/<name>.json.

The kernels are table-driven; every table below is walked with a
wavefront-uniform index, so the compiler reads it through the scalar cache.

  rx_s[NREACT*RX_MAX]   reactant species of each reaction, repeated by
                        multiplicity, -1 padded (index >= NVAR: fixed species);
                        A(r) = RCT(r) * prod conc(rx_s)    (Fun_SPLIT's A(r))
  fv_ptr/fv_r/fv_c      Vdot(i) = sum_t fv_c[t] * A(fv_r[t])  (net stoichiometry)
  pt_r/pt_m/pt_s        partial derivative b: d A(pt_r)/d V(j) =
                        pt_m * RCT(pt_r) * prod conc(pt_s[b][0..RX_MAX-2])
  jv_ptr/jv_b/jv_c      JVS(k) = sum_t jv_c[t] * partial(jv_b[t])  (Jac_SP)
  lu_crow/lu_diag/lu_icol   KPP's sparse LU pattern (0-based)
  le_*, lp_*            the KppDecomp elimination as an op stream: for each
                        lower entry e (row-major): a = JVS(le_kk)/JVS(le_dj);
                        JVS(le_kk)=a; then JVS(lp_dst) -= a*JVS(lp_src) for
                        lp in [le_p0, le_p1)
  cu_*, sf_*, sb_*      the same sums as flat segment lists (see build_tables)
                        for the persistent solver's batched execution
  wv_*, WV_*            the cell-per-wave solver's programs, LDS and scratch
                        layouts (wave_sched.py)
  lc<S>_lu, lc<S>_sv    the persistent solver's operand-cache schedules with
                        S slots per lane (lane_cache.py)
Run:  python emit_tables.py <mech> [<mech> ...]
"""
import json
import sys
from pathlib import Path

sys.path.insert(0, str(Path(__file__).resolve().parent))
import lane_cache  # noqa: E402
import wave_sched  # noqa: E402

HERE = Path(__file__).resolve().parent
RX_MAX = 3


def arr(ctype, name, vals, per_line=16, fmt=str, align=0):
    vals = list(vals)
    if not vals:
        vals = [0]
    lines = []
    for i in range(0, len(vals), per_line):
        lines.append("  " + ", ".join(fmt(v) for v in vals[i:i + per_line]) + ",")
    # internal linkage: the header is included by more than one translation unit
    al = " __attribute__((aligned(%d)))" % align if align else ""
    return "static __constant__ %s %s[%d]%s = {\n%s\n};\n" % (ctype, name, len(vals), al,
                                                               "\n".join(lines))


def dfmt(v):
    r = repr(float(v))
    return r if ("e" in r or "." in r or "inf" in r or "nan" in r) else r + ".0"


def build_tables(m):
    nvar, nfix, nreact = m["nvar"], m["nfix"], m["nreact"]
    rx = m["reactions"]
    # reactants expanded by multiplicity
    rx_s = []
    for r in rx:
        lst = []
        for s, k in r["reactants"]:
            lst += [s] * k
        assert len(lst) <= RX_MAX
        rx_s += lst + [-1] * (RX_MAX - len(lst))
    # Vdot gather
    terms = [[] for _ in range(nvar)]
    for ri, r in enumerate(rx):
        for s, c in r["stoich"]:
            terms[s].append((ri, c))
    fv_ptr, fv_r, fv_c = [0], [], []
    for i in range(nvar):
        for ri, c in sorted(terms[i]):
            fv_r.append(ri)
            fv_c.append(c)
        fv_ptr.append(len(fv_r))
    # partials
    pt_r, pt_m, pt_s, pidx = [], [], [], {}
    for ri, r in enumerate(rx):
        for s, k in r["reactants"]:
            if s >= nvar:
                continue
            others = []
            for t, kt in r["reactants"]:
                others += [t] * (kt - 1 if t == s else kt)
            pidx[(ri, s)] = len(pt_r)
            pt_r.append(ri)
            pt_m.append(float(k))
            pt_s += others + [-1] * (RX_MAX - 1 - len(others))
    # JVS gather in LU layout
    lu_irow, lu_icol = m["lu_irow"], m["lu_icol"]
    # a mechanism without a sparse Jacobian (carbon: forward Euler) has none
    nz = m["lu_nonzero"] if lu_irow else 0
    pos = {(lu_irow[k], lu_icol[k]): k for k in range(nz)}
    jterms = [[] for _ in range(nz)]
    for ri, r in enumerate(rx if nz else []):
        for s, k in r["reactants"]:
            if s >= nvar:
                continue
            b = pidx[(ri, s)]
            for i, c in r["stoich"]:
                kk = pos.get((i, s))
                if kk is None:
                    raise ValueError("Jacobian entry (%d,%d) outside LU pattern" % (i, s))
                jterms[kk].append((b, c))
    jv_ptr, jv_b, jv_c = [0], [], []
    for k in range(nz):
        for b, c in sorted(jterms[k]):
            jv_b.append(b)
            jv_c.append(c)
        jv_ptr.append(len(jv_b))
    # elimination op stream (row-wise Doolittle, KppDecomp order)
    crow, diag = m["lu_crow"], m["lu_diag"]
    le_kk, le_dj, le_p0, le_p1, le_row = [], [], [], [], []
    lp_dst, lp_src = [], []
    row_e0 = []
    for k in range(nvar if crow else 0):
        row_e0.append(len(le_kk))
        rowpos = {lu_icol[q]: q for q in range(crow[k], crow[k + 1])}
        for kk in range(crow[k], diag[k]):
            j = lu_icol[kk]
            le_kk.append(kk)
            le_dj.append(diag[j])
            le_row.append(k)
            le_p0.append(len(lp_dst))
            for jj in range(diag[j] + 1, crow[j + 1]):
                lp_dst.append(rowpos[lu_icol[jj]])
                lp_src.append(jj)
            le_p1.append(len(lp_dst))
    row_e0.append(len(le_kk))
    # the same elimination as dot products per LU position (row-Crout order):
    # JVS(q) for q=(k,c) receives -L(k,j)*U(j,c) for the pivots j < min(c,k)
    # of row k, in increasing j (the order KppDecomp applies them), then, for
    # c < k, is divided by the pivot U(c,c).
    cr_ptr, cr_l, cr_u, cr_div = [0], [], [], []
    for k in range(nvar if crow else 0):
        rowpos = {lu_icol[q]: q for q in range(crow[k], crow[k + 1])}
        lower = [lu_icol[q] for q in range(crow[k], diag[k])]
        for q in range(crow[k], crow[k + 1]):
            c = lu_icol[q]
            for j in lower:
                if j >= min(c, k):
                    break
                # U(j,c) in row j's pattern?
                for jj in range(diag[j] + 1, crow[j + 1]):
                    if lu_icol[jj] == c:
                        cr_l.append(rowpos[j])
                        cr_u.append(jj)
                        break
            cr_ptr.append(len(cr_l))
            cr_div.append(diag[c] if c < k else -1)
    assert len(cr_l) == len(lp_dst)
    # flat segmented forms for latency-batched execution: each segment's
    # first term carries its initial value with l index LU_NONZERO, the
    # solver's constant entry -lu[LU_NONZERO] = 1 (see the persistent slot)
    one = m["lu_nonzero"]
    #  cu_*: the U part of each Crout row (positions q >= diag(k)), which are
    #        independent once the row's L part is done: segment = position
    cu_row, cu_sp, cu_q, cu_l, cu_u = [], [0], [], [], []
    for k in range(nvar if crow else 0):
        cu_row.append(len(cu_q))
        for q in range(diag[k], crow[k + 1]):
            cu_l.append(one)
            cu_u.append(q)
            for t in range(cr_ptr[q], cr_ptr[q + 1]):
                cu_l.append(cr_l[t])
                cu_u.append(cr_u[t])
            cu_q.append(q)
            cu_sp.append(len(cu_l))
    cu_row.append(len(cu_q))
    #  sf_*/sb_*: forward / backward substitution, segment = row, in
    #  batches of consecutive rows that do not read each other's results
    def solve_tables(order, lower):
        sp, tl, tx, batch = [0], [], [], [0]
        cur = set()
        for n_, i in enumerate(order):
            lo, hi = (crow[i], diag[i]) if lower else (diag[i] + 1, crow[i + 1])
            deps = {lu_icol[k] for k in range(lo, hi)}
            if cur & deps:
                batch.append(n_)
                cur = set()
            cur.add(i)
            tl.append(one)
            tx.append(i)
            for k in range(lo, hi):
                tl.append(k)
                tx.append(lu_icol[k])
            sp.append(len(tl))
        batch.append(len(order))
        return sp, tl, tx, batch
    if crow:
        sf_sp, sf_l, sf_x, sf_batch = solve_tables(list(range(nvar)), True)
        sb_sp, sb_l, sb_x, sb_batch = solve_tables(list(range(nvar - 1, -1, -1)), False)
    else:
        sf_sp = sf_l = sf_x = sf_batch = sb_sp = sb_l = sb_x = sb_batch = []
    return dict(cu_row=cu_row, cu_sp=cu_sp, cu_q=cu_q, cu_l=cu_l, cu_u=cu_u,
                sf_sp=sf_sp, sf_l=sf_l, sf_x=sf_x, sf_batch=sf_batch,
                sb_sp=sb_sp, sb_l=sb_l, sb_x=sb_x, sb_batch=sb_batch,
                rx_s=rx_s, fv_ptr=fv_ptr, fv_r=fv_r, fv_c=fv_c, pt_r=pt_r, pt_m=pt_m,
                pt_s=pt_s, jv_ptr=jv_ptr, jv_b=jv_b, jv_c=jv_c, le_kk=le_kk, le_dj=le_dj,
                le_p0=le_p0, le_p1=le_p1, le_row=le_row, lp_dst=lp_dst, lp_src=lp_src,
                row_e0=row_e0, cr_ptr=cr_ptr, cr_l=cr_l, cr_u=cr_u, cr_div=cr_div)


def emit(name):
    m = json.loads((HERE.parent / "mech" / ("%s.json" % name)).read_text())
    t = build_tables(m)
    out = ["/* Generated by geos-chem_amd/mechgen/emit_tables.py from mech/%s.json"
           " -- do not edit. */" % name, "#pragma once", "",
           "namespace mech_%s {" % name,
           "constexpr int NVAR = %d, NFIX = %d, NSPEC = %d, NREACT = %d;"
           % (m["nvar"], m["nfix"], m["nspec"], m["nreact"]),
           "constexpr int LU_NZ = %d, NPART = %d, RX_MAX = %d;" % (m["lu_nonzero"], len(t["pt_r"]), RX_MAX),
           "constexpr int NFV = %d, NJV = %d, NLE = %d, NLP = %d;"
           % (len(t["fv_r"]), len(t["jv_b"]), len(t["le_kk"]), len(t["lp_dst"])),
           "constexpr int NSF_BATCH = %d, NSB_BATCH = %d;"
           % (max(0, len(t["sf_batch"]) - 1), max(0, len(t["sb_batch"]) - 1)), ""]
    ints = ["rx_s", "fv_ptr", "fv_r", "pt_r", "pt_s", "jv_ptr", "jv_b", "le_kk", "le_dj",
            "le_p0", "le_p1", "le_row", "lp_dst", "lp_src", "row_e0", "cr_ptr", "cr_l", "cr_u",
            "cr_div", "cu_row", "cu_sp", "cu_q", "cu_l", "cu_u", "sf_sp", "sf_l", "sf_x",
            "sf_batch", "sb_sp", "sb_l", "sb_x", "sb_batch"]
    for k in ints:
        out.append(arr("int", k, t[k]))
    for k in ("fv_c", "pt_m", "jv_c"):
        out.append(arr("double", k, t[k], per_line=6, fmt=dfmt))
    for k, key in (("lu_crow", "lu_crow"), ("lu_diag", "lu_diag"), ("lu_icol", "lu_icol")):
        out.append(arr("int", k, m[key][: (m["nvar"] + 1) if k != "lu_icol" else None]))
    # cell-per-wave solver: LDS layout, global scratch layout and programs
    prog = wave_sched.build_wave_program(m, t)
    L = prog["layout"]
    out.append("constexpr int WV_ZERO = %d, WV_ONE = %d, WV_X0 = %d, WV_Y0 = %d, WV_SINK = %d, "
               "WV_SH = %d;" % (L["ZERO"], L["ONE"], L["X0"], L["Y0"], L["SINK"], L["SH_SIZE"]))
    out.append("constexpr int WV_G_AB = %d, WV_NAB = %d, WV_G_F0 = %d, WV_G_K = %d, WV_G_SIZE = %d, "
               "WV_NCOEF = %d;" % (L["G_AB"], L["NAB"], L["G_F0"], L["G_K"], L["G_SIZE"],
                                   len(prog["coef"])))
    arrays, consts = wave_sched.emit_arrays(prog)
    have = {P for P in ("fv", "jv", "lu", "sf", "sb") if P in prog}
    cmap = dict(consts)
    for P in ("fv", "jv", "lu", "sf", "sb"):
        out.append("constexpr int WV_%s_NB = %d, WV_%s_NR = %d;"
                   % (P.upper(), cmap.get("WV_%s_NB" % P.upper(), 0), P.upper(),
                      cmap.get("WV_%s_NR" % P.upper(), 0)))
    for aname, ctype, vals in arrays:   # program words are read as 16-byte blocks
        out.append(arr(ctype, aname, vals, per_line=6 if ctype == "double" else 16,
                       fmt=dfmt if ctype == "double" else str,
                       align=16 if aname.endswith("_ops") else 0))
    for P in ("fv", "jv", "lu", "sf", "sb"):   # empty programs of a mechanism without an LU
        if P not in have:
            for suf in ("ops", "flag"):
                out.append(arr("unsigned", "wv_%s_%s" % (P, suf), [0, 0, 0, 0],
                               align=16 if suf == "ops" else 0))
    # persistent lane solver: operand-cache schedules of the LU and
    # substitution sections (lane_cache.py), 16 pad words at the end so a
    # record's words can be loaded whole
    for S in lane_cache.SLOTS:
        lw, sw = lane_cache.build(m, t, S)
        out.append(arr("unsigned", "lc%d_lu" % S, lw + [0] * 16))
        out.append(arr("unsigned", "lc%d_sv" % S, sw + [0] * 16))
    out.append("}  // namespace mech_%s" % name)
    gen = HERE.parent / "csrc" / "gen"
    gen.mkdir(exist_ok=True)
    (gen / ("mech_%s.h" % name)).write_text("\n".join(out) + "\n")
    print("%s: NFV=%d NPART=%d NJV=%d NLE=%d NLP=%d" % (name, len(t["fv_r"]), len(t["pt_r"]),
                                                       len(t["jv_b"]), len(t["le_kk"]), len(t["lp_dst"])))


if __name__ == "__main__":
    for n in sys.argv[1:] or ["fullchem", "Hg", "carbon"]:
        emit(n)
