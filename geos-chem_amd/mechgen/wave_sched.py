"""Mechanism compiler, wave programs: the schedules of the cell-per-wave
Rodas3 solver (csrc/wave_solver.hip), built from the tables of
emit_tables.build_tables.

In that solver one wavefront (64 lanes) integrates one cell whose state
lives in LDS.  Every sum of the solve -- a Vdot row, a Jacobian entry, an LU
position (row-Crout form), a row of the forward or backward substitution --
is computed by one lane, term by term in the table order, so each value is
rounded exactly as in the lane-per-cell solvers.  A sum whose terms become
available at different times (an LU position needs L(k,j) and U(j,c) of
earlier positions; a substitution row needs x of earlier rows) is split into
chunks at dependency levels: the chunk of level l holds the next terms whose
operands are final before level l; the partial sum is kept in the sum's own
LDS slot between chunks.  Chunks of one level are independent; they are
sorted by length and dealt to the 64 lanes in rounds.

Encoding (all LDS addresses are element indices of the solver's LDS block,
whose layout `layout()` fixes):
  <P>_lev[NLEV+1]   first round of each level (levels are separated by a
                    wave-wide LDS ordering point)
  <P>_rs[NR+1]      first step of each round (steps of a round = its longest
                    chunk; shorter chunks are padded with exact no-op terms)
  <P>_head[(NR+1)*64] per round and lane: low 16 bits the destination slot
                    (0xFFFF: lane idle), high 16 bits the divisor slot + 1
                    (LU and backward substitution: the final chunk divides;
                    Jacobian: 1 if the slot is a diagonal entry); one idle
                    round ends the array (the solver loads a round ahead)
  <P>_term[(NS+8)*64] per step and lane: low 16 bits operand a, high 16 bits
                    operand b; the step is acc = fma(b, a, acc) for Vdot and
                    Jacobian (b a coefficient slot) and acc = fma(-a, b, acc)
                    for LU and substitution.  A pad term has a = ZERO and b
                    = the -0.0 coefficient (Vdot, Jacobian) or ZERO: its
                    product is -0.0, which leaves every acc unchanged.
                    Eight steps of pads end the array (the solver loads
                    eight steps at a time).
Phases P: fv (Vdot), jv (Jacobian in LU layout), lu (row-Crout), sf
(forward substitution), sb (backward substitution).
"""
import numpy as np

LANES = 64
IDLE = 0xFFFF


def layout(m, t, ncoef):
    """LDS element offsets of the cell-per-wave solver for one mechanism."""
    nvar, nspec, nreact = m["nvar"], m["nspec"], m["nreact"]
    lu_nz = m["lu_nonzero"] if m["lu_irow"] else 1
    scr = max(nreact, len(t["pt_r"]))
    L = {"LU0": 0, "ZERO": lu_nz}
    L["X0"] = lu_nz + 1
    L["Y0"] = L["X0"] + nvar
    L["YN0"] = L["Y0"] + nspec
    L["AB0"] = L["YN0"] + nvar
    L["COEF0"] = L["AB0"] + scr
    L["SH_SIZE"] = L["COEF0"] + ncoef
    assert L["SH_SIZE"] < IDLE
    return L


def _rounds(levels):
    """levels: list (per level) of chunks (dest, hi, [(a, b), ...]).
    Returns lev, rs, head, term arrays (interleaved by lane)."""
    lev, rs, head, term = [0], [0], [], []
    for chunks in levels:
        ch = sorted(chunks, key=lambda c: -len(c[2]))
        for r0 in range(0, len(ch), LANES):
            grp = ch[r0:r0 + LANES]
            ns = max(len(c[2]) for c in grp)
            h = [IDLE] * LANES
            tt = [[None] * LANES for _ in range(ns)]
            for ln, (dest, hi, terms) in enumerate(grp):
                h[ln] = dest | (hi << 16)
                for s, ab in enumerate(terms):
                    tt[s][ln] = ab
            head += h
            term += [row for row in tt]
            rs.append(rs[-1] + ns)
        lev.append(len(rs) - 1)
    return lev, rs, head, term


BLOCK = 8   # steps the solver loads at once; the term array ends with one block of pads


def _encode_terms(term_rows, pad):
    out = []
    for row in term_rows:
        for ab in row:
            a, b = ab if ab is not None else pad
            out.append(a | (b << 16))
    return out


def _encode_terms_padded(term_rows, pad):
    """Term words plus BLOCK steps of pads, so the solver may load a whole
    block past any round's last step."""
    return _encode_terms(term_rows, pad) + [pad[0] | (pad[1] << 16)] * (BLOCK * LANES)


def _split(terms_ready, final_extra_level):
    """Chunks of one sum: terms_ready[i] = earliest level of term i; the sum
    goes in order, so term i runs at max(its level, previous term's level).
    final_extra_level: earliest level of the final operation (division), or
    None.  Returns [(level, [term indices], is_final)] and the final level."""
    chunks = []
    lev = 0
    cur = None
    for i, r in enumerate(terms_ready):
        lv = max(lev, r)
        if cur is None or lv != cur[0]:
            cur = (lv, [])
            chunks.append(cur)
        cur[1].append(i)
        lev = lv
    if final_extra_level is not None:
        lv = max(lev, final_extra_level)
        if cur is None or lv != cur[0]:
            cur = (lv, [])
            chunks.append(cur)
        lev = lv
    return chunks, lev


def build_wave_program(m, t):
    """Wave-program tables and LDS layout for one mechanism."""
    nvar = m["nvar"]
    crow, diag, icol = m["lu_crow"], m["lu_diag"], m["lu_icol"]
    has_lu = bool(m["lu_irow"])
    coefs = [-0.0]
    cidx = {}
    for c in list(t["fv_c"]) + list(t["jv_c"]):
        key = float(c)
        if key not in cidx:
            cidx[key] = len(coefs)
            coefs.append(key)
    L = layout(m, t, len(coefs))
    ZERO, X0, AB0, C0 = L["ZERO"], L["X0"], L["AB0"], L["COEF0"]
    out = {"layout": L, "coef": coefs}

    # Vdot: one chunk per row, all in level 0
    fv_ptr, fv_r, fv_c = t["fv_ptr"], t["fv_r"], t["fv_c"]
    ch = []
    for i in range(nvar):
        ch.append((X0 + i, 0, [(AB0 + fv_r[k], C0 + cidx[float(fv_c[k])])
                               for k in range(fv_ptr[i], fv_ptr[i + 1])]))
    out["fv"] = _rounds([ch]), (ZERO, C0)

    if not has_lu:
        return out
    nz = m["lu_nonzero"]
    # Jacobian entries in LU layout (sign and diagonal applied at the store)
    jv_ptr, jv_b, jv_c = t["jv_ptr"], t["jv_b"], t["jv_c"]
    isdiag = set(diag[:nvar])
    ch = []
    for q in range(nz):
        ch.append((q, 1 if q in isdiag else 0, [(AB0 + jv_b[k], C0 + cidx[float(jv_c[k])])
                                                for k in range(jv_ptr[q], jv_ptr[q + 1])]))
    out["jv"] = _rounds([ch]), (ZERO, C0)

    # row-Crout LU: position q = (k, c) sums -L(k,j)*U(j,c) in cr order, then
    # divides by U(c,c) when c < k
    cr_ptr, cr_l, cr_u, cr_div = t["cr_ptr"], t["cr_l"], t["cr_u"], t["cr_div"]
    fin = [0] * nz
    levels = {}
    for q in range(nz):
        rng = range(cr_ptr[q], cr_ptr[q + 1])
        ready = [max(fin[cr_l[k]], fin[cr_u[k]]) + 1 for k in rng]
        dv = cr_div[q]
        chunks, fl = _split(ready, fin[dv] + 1 if dv >= 0 else None)
        for n_, (lv, idx) in enumerate(chunks):
            last = n_ == len(chunks) - 1
            hi = (dv + 1) if (last and dv >= 0) else 0
            levels.setdefault(lv, []).append(
                (q, hi, [(cr_l[cr_ptr[q] + i], cr_u[cr_ptr[q] + i]) for i in idx]))
        fin[q] = fl
    out["lu"] = _rounds([levels[l] for l in sorted(levels)]), (ZERO, ZERO)
    out["lu_fin"] = fin

    # substitutions: x(i) = x(i) - sum L(i,c) x(c), then backward
    # x(i) = (x(i) - sum U(i,c) x(c)) / U(i,i), terms in LU order
    def subst(order, lower):
        fx = [0] * nvar
        lv_map = {}
        for i in order:
            lo, hi = (crow[i], diag[i]) if lower else (diag[i] + 1, crow[i + 1])
            ready = [fx[icol[k]] + 1 for k in range(lo, hi)]
            chunks, fl = _split(ready, None if lower else 1)
            for n_, (lv, idx) in enumerate(chunks):
                last = n_ == len(chunks) - 1
                dvh = (diag[i] + 1) if (last and not lower) else 0
                lv_map.setdefault(lv, []).append(
                    (X0 + i, dvh, [(lo + j, X0 + icol[lo + j]) for j in idx]))
            fx[i] = fl
        return _rounds([lv_map[l] for l in sorted(lv_map)])
    out["sf"] = subst(range(nvar), True), (ZERO, ZERO)
    out["sb"] = subst(range(nvar - 1, -1, -1), False), (ZERO, ZERO)
    return out


def emit_arrays(prog):
    """(name, ctype, values) of the wave-program tables."""
    arrays = []
    for P in ("fv", "jv", "lu", "sf", "sb"):
        if P not in prog:
            continue
        (lev, rs, head, term_rows), pad = prog[P]
        arrays.append(("wv_%s_lev" % P, "int", lev))
        arrays.append(("wv_%s_rs" % P, "int", rs))
        arrays.append(("wv_%s_head" % P, "unsigned", head + [IDLE] * LANES))
        arrays.append(("wv_%s_term" % P, "unsigned", _encode_terms_padded(term_rows, pad)))
    arrays.append(("wv_coef", "double", prog["coef"]))
    return arrays


# ---------------------------------------------------------------------------
# Emulation (tests): run a phase's program on a flat LDS image with a given
# arithmetic for one step, to show it performs the same operations in the
# same order as the sequential table walk.
def run_phase(prog, P, sh, step, init_from_dest, store):
    """Executes phase P on the array sh (modified in place).  step(acc, a, b)
    -> acc; init_from_dest: acc starts at sh[dest] (else +0.0); store(sh,
    dest, hi, acc, final_hi) writes the result."""
    (lev, rs, head, term_rows), pad = prog[P]
    for l in range(len(lev) - 1):
        writes = []
        for r in range(lev[l], lev[l + 1]):
            for ln in range(LANES):
                h = head[r * LANES + ln]
                dest = h & 0xFFFF
                if dest == IDLE:
                    continue
                acc = sh[dest] if init_from_dest else 0.0
                for s in range(rs[r], rs[r + 1]):
                    ab = term_rows[s][ln]
                    a, b = ab if ab is not None else pad
                    acc = step(acc, sh[a], sh[b])
                writes.append((dest, h >> 16, acc))
        # a level's results become visible only after the level
        for dest, hi, acc in writes:
            store(sh, dest, hi, acc)
    return sh


def level_dependencies_ok(prog, P):
    """Every slot a chunk reads (operands, divisor) that the phase writes is
    final (its last chunk done) before the chunk's level; no two chunks of
    one level write the same slot."""
    (lev, rs, head, term_rows), pad = prog[P]
    last = {}
    for l in range(len(lev) - 1):
        for r in range(lev[l], lev[l + 1]):
            for ln in range(LANES):
                d = head[r * LANES + ln] & 0xFFFF
                if d != IDLE:
                    last[d] = l
    for l in range(len(lev) - 1):
        dests = set()
        for r in range(lev[l], lev[l + 1]):
            for ln in range(LANES):
                h = head[r * LANES + ln]
                d = h & 0xFFFF
                if d == IDLE:
                    continue
                if d in dests:
                    return False
                dests.add(d)
                reads = []
                if P in ("lu", "sb") and (h >> 16):
                    reads.append((h >> 16) - 1)
                for s in range(rs[r], rs[r + 1]):
                    ab = term_rows[s][ln]
                    if ab is not None:
                        reads += [ab[0], ab[1]]
                for o in reads:
                    if o in last and last[o] >= l:
                        return False
    return True
