"""Mechanism compiler, wave programs: the schedules of the cell-per-wave
Rodas3 solver (csrc/wave_solver.hip), built from the tables of
emit_tables.build_tables.

In that solver one wavefront (64 lanes) integrates one cell whose LU,
solve vector and state live in LDS.  Every sum of the solve -- a Vdot row,
a Jacobian entry, an LU position (row-Crout form), a row of the forward or
backward substitution -- is computed by one lane, term by term in the table
order, so each value is rounded exactly as in the lane-per-cell solvers.  A
sum whose terms become available at different times (an LU position needs
L(k,j) and U(j,c) of earlier positions; a substitution row needs x of
earlier rows) is split into chunks at dependency levels: the chunk of level
l holds the next terms whose operands are final before level l; the partial
sum is kept in the sum's own LDS slot between chunks.  Chunks of one level
are independent; they are sorted by length and dealt to the 64 lanes in
rounds.  A round's steps are padded to whole blocks of BSTEP steps.

Operand spaces.  fv (Vdot) and jv (Jacobian): operand a is an element of
the wave's reaction scratch in global memory (A(r) or the partials;
element NAB is a +0.0 pad slot), operand b an element of the mechanism's
coefficient table (element 0 is -0.0); a step is acc = fma(b, a, acc) from
acc = +0.0.  lu, sf, sb: both operands are LDS elements; a step is
acc = fma(-a, b, acc) from acc = the destination slot.

Encoding (the solver reads BYTE offsets, element index * 8):
  wv_<P>_ops   [NB + PF + 4][2][64] uint4 words: block k, half h, lane: the
               lane's ops 2h and 2h+1 of the block, (a, b, a', b').  A
               round is its header op followed by its terms, padded with
               pad ops to whole blocks of BSTEP ops.  Header op: LU and
               substitutions (dest, divisor) -- the partial sum starts at
               LDS[dest], the result is divided by LDS[divisor], the +1.0
               slot when the lane does not divide; Vdot and Jacobian
               (dest, diagonal flag) with the sum starting at +0.0.  An idle
               lane's header names its own sink slot (SINK + lane).  Term
               op (a, b): LDS byte offsets (LU, substitutions), or the
               reaction-scratch and coefficient byte offsets (Vdot,
               Jacobian).  Pad op: (ZERO, ZERO) in LDS phases and (NAB, 0)
               in Vdot/Jacobian -- products -0.0 that leave every sum
               unchanged.  In the LDS phases a one-block round of sink
               headers and pad ops follows every level but the last: the
               solver reads a block's operands while the block before it
               runs, so the first block of a level must not follow the
               level's last block directly.  NB is padded to a multiple of
               PF (the first padding block starts a round of sink headers:
               the solver stores a lane's value after every block); PF + 4
               pad blocks follow (the solver loads program words PF blocks
               and operands up to 3 blocks ahead).
  wv_<P>_flag  4 bits per block (block k in bits 4(k%8)..4(k%8)+3 of word
               k/8): 1 first block of a round, 2 last block of a round,
               4 the round ends a level, 8 some lane of the round divides
               (LU, backward substitution).
  WV_<P>_NB    blocks (a multiple of PF), WV_<P>_NR rounds (with the
               levels' sink rounds).
Phases P: fv (Vdot), jv (Jacobian in LU layout), lu (row-Crout), sf
(forward substitution), sb (backward substitution).
"""
import numpy as np

LANES = 64
IDLE = 0xFFFF


BSTEP = 4    # ops per block (two 16-byte loads of program words per lane)
PF = 8       # blocks of program words the solver loads ahead; NB is a multiple of it
XPAD = 4     # pad blocks past NB + PF (operands are loaded up to 3 blocks ahead)
F_START, F_END, F_LEVEL, F_DIV = 1, 2, 4, 8


def layout(m, t):
    """Element offsets of the cell-per-wave solver's LDS block (LU, a +0.0
    slot, a +1.0 slot, the solve vector X -- also the stage point Ynew and
    the scaled error terms --, the state Y, one sink slot per lane) and of
    a wave's global scratch (rate constants, reaction products / partials,
    a +0.0 slot, Fcn0, the four stage vectors K)."""
    nvar, nspec, nreact = m["nvar"], m["nspec"], m["nreact"]
    lu_nz = m["lu_nonzero"] if m["lu_irow"] else 1
    nab = max(nreact, len(t["pt_r"]))
    L = {"LU0": 0, "ZERO": lu_nz, "ONE": lu_nz + 1}
    L["X0"] = lu_nz + 2
    L["Y0"] = L["X0"] + nvar
    L["SINK"] = L["Y0"] + nspec
    L["SH_SIZE"] = L["SINK"] + LANES
    assert 8 * L["SH_SIZE"] < IDLE
    L["G_RCT"] = 0
    L["G_AB"] = nreact
    L["NAB"] = nab
    L["G_F0"] = nreact + nab + 1
    L["G_K"] = L["G_F0"] + nvar
    L["G_SIZE"] = L["G_K"] + 4 * nvar
    assert 8 * (nab + 1) < IDLE
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
    """Wave-program tables and LDS layout for one mechanism.  Operands are
    element indices of their space (see the module docstring)."""
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
    L = layout(m, t)
    ZERO, X0, NAB = L["ZERO"], L["X0"], L["NAB"]
    out = {"layout": L, "coef": coefs}

    # Vdot: one chunk per row, all in level 0; operands (reaction product,
    # coefficient)
    fv_ptr, fv_r, fv_c = t["fv_ptr"], t["fv_r"], t["fv_c"]
    ch = []
    for i in range(nvar):
        ch.append((X0 + i, 0, [(fv_r[k], cidx[float(fv_c[k])])
                               for k in range(fv_ptr[i], fv_ptr[i + 1])]))
    out["fv"] = _rounds([ch]), (NAB, 0)

    if not has_lu:
        return out
    nz = m["lu_nonzero"]
    # Jacobian entries in LU layout (sign and diagonal applied at the store);
    # operands (partial, coefficient)
    jv_ptr, jv_b, jv_c = t["jv_ptr"], t["jv_b"], t["jv_c"]
    isdiag = set(diag[:nvar])
    ch = []
    for q in range(nz):
        ch.append((q, 1 if q in isdiag else 0, [(jv_b[k], cidx[float(jv_c[k])])
                                                for k in range(jv_ptr[q], jv_ptr[q + 1])]))
    out["jv"] = _rounds([ch]), (NAB, 0)

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


def encode(prog, P):
    """The solver's arrays of phase P (see the module docstring): ops words
    [NB + PF + XPAD][2][64][4], flag words, NB, NR."""
    (lev, rs, head, term_rows), pad = prog[P]
    L = prog["layout"]
    nr = len(rs) - 1
    lds = P not in ("fv", "jv")
    lev_end = set(lev[1:])
    blocks, flags = [], []
    npad = 0
    padop = (8 * L["ZERO"], 8 * L["ZERO"]) if lds else (8 * pad[0], 8 * pad[1])

    def hdr(r, ln):
        h = head[r * LANES + ln]
        d, hi = h & 0xFFFF, h >> 16
        if d == IDLE:
            return (8 * (L["SINK"] + ln), 8 * L["ONE"] if lds else 0)
        if not lds:
            return (8 * d, hi)
        if P in ("lu", "sb") and hi:
            return (8 * d, 8 * (hi - 1))
        return (8 * d, 8 * L["ONE"])

    for r in range(nr):
        rows = term_rows[rs[r]:rs[r + 1]]
        nops = 1 + len(rows)
        nblk = -(-nops // BSTEP)
        lanes = []
        divides = False
        for ln in range(LANES):
            ops = [hdr(r, ln)]
            for row in rows:
                ab = row[ln]
                ops.append((8 * ab[0], 8 * ab[1]) if ab is not None else padop)
            ops += [padop] * (nblk * BSTEP - len(ops))
            lanes.append(ops)
            h = head[r * LANES + ln]
            if P in ("lu", "sb") and (h & 0xFFFF) != IDLE and (h >> 16):
                divides = True
        for bk in range(nblk):
            blocks.append([lanes[ln][bk * BSTEP:(bk + 1) * BSTEP] for ln in range(LANES)])
            f = (F_START if bk == 0 else 0) | (F_END if bk == nblk - 1 else 0)
            if bk == nblk - 1 and (r + 1) in lev_end:
                f |= F_LEVEL
            if bk == nblk - 1 and divides:
                f |= F_DIV
            flags.append(f)
        if lds and (r + 1) in lev_end and r + 1 < nr:
            # the solver requests a block's LDS operands while the block
            # before it runs, before that block's stores: a round of sink
            # headers and pad ops after a level's last block keeps the next
            # level's first operands clear of the level's last results
            blocks.append([[(8 * (L["SINK"] + ln), 8 * L["ONE"])] + [padop] * (BSTEP - 1)
                           for ln in range(LANES)])
            flags.append(F_START | F_END)
            npad += 1
    nb = len(blocks)
    nb_pad = -(-nb // PF) * PF
    empty = [[padop] * BSTEP for _ in range(LANES)]
    blocks += [empty] * (nb_pad - nb + PF + XPAD)
    flags += [0] * (nb_pad - nb + PF + XPAD)
    if nb_pad > nb:
        # the solver stores a lane's value after every block: the blocks
        # that pad the last round's end to PF start a round of sink
        # headers, so they store into the sinks
        blocks[nb] = [[(8 * (L["SINK"] + ln), 8 * L["ONE"] if lds else 0)] + [padop] * (BSTEP - 1)
                      for ln in range(LANES)]
        flags[nb] = F_START
    flags += [0] * (-len(flags) % 8)
    ops = []
    for blk in blocks:
        for h in range(2):
            for ln in range(LANES):
                o0, o1 = blk[ln][2 * h], blk[ln][2 * h + 1]
                ops += [o0[0], o0[1], o1[0], o1[1]]
    flag = []
    for k in range(0, len(flags), 8):
        w = 0
        for i, f in enumerate(flags[k:k + 8]):
            w |= f << (4 * i)
        flag.append(w)
    return ops, flag, nb_pad, nr + npad


def emit_arrays(prog):
    """(name, ctype, values) of the wave-program tables, and the constants
    (name, value)."""
    arrays, consts = [], []
    for P in ("fv", "jv", "lu", "sf", "sb"):
        if P not in prog:
            continue
        ops, flag, nb, nr = encode(prog, P)
        arrays.append(("wv_%s_ops" % P, "unsigned", ops))
        arrays.append(("wv_%s_flag" % P, "unsigned", flag))
        consts.append(("WV_%s_NB" % P.upper(), nb))
        consts.append(("WV_%s_NR" % P.upper(), nr))
    arrays.append(("wv_coef", "double", prog["coef"]))
    return arrays, consts


# ---------------------------------------------------------------------------
# Emulation (tests): run a phase's program on a flat LDS image with a given
# arithmetic for one step, to show it performs the same operations in the
# same order as the sequential table walk.
def run_phase(prog, P, sh, step, init_from_dest, store, ab=None, coef=None):
    """Executes phase P from its encoded arrays in the solver's order
    (csrc/wave_solver.hip, run_prog): block by block, the operands of a
    block read D blocks before it runs, before the stores of the blocks in
    between (D = 3 for Vdot and Jacobian, 1 otherwise), a round's header op
    naming the start of its sum, and every lane's value stored after every
    block.  sh: the LDS block (element list, modified in place); ab, coef:
    the reaction scratch (with its +0.0 pad slot) and the coefficient table
    (fv, jv operands).  step(acc, a, b) -> acc; init_from_dest: acc starts
    at sh[dest] (else +0.0); store(sh, dest, hi, acc) writes a value
    (element indices; hi: divisor element + 1 when the round's last block
    divides and the divisor is not the +1.0 slot, the Jacobian's diagonal
    flag, else 0)."""
    ops, flag, nb, nr = encode(prog, P)
    L = prog["layout"]
    lds = P not in ("fv", "jv")
    D = 1 if lds else 3
    NAB = L["NAB"]

    def flag_of(k):
        return (flag[k // 8] >> (4 * (k % 8))) & 15

    def fetch(k):
        f = flag_of(k)
        out = []
        for ln in range(LANES):
            w = [ops[((2 * k + h) * LANES + ln) * 4 + j] for h in range(2) for j in range(4)]
            u = [(w[0], w[1]), (w[2], w[3]), (w[4], w[5]), (w[6], w[7])]
            vals = []
            for q, (a, b) in enumerate(u):
                if lds:
                    vals.append((sh[a // 8], sh[b // 8]))
                elif q == 0 and f & F_START:
                    vals.append((ab[NAB], coef[0]))
                else:
                    vals.append((ab[a // 8], coef[b // 8]))
            out.append((u[0], vals))
        return out

    pending = {k: fetch(k) for k in range(D)}
    acc = [0.0] * LANES
    cur = [(0, 0)] * LANES
    dv = [1.0] * LANES
    r = 0
    for k in range(nb):
        f = flag_of(k)
        pending[k + D] = fetch(k + D)
        blk = pending.pop(k)
        for ln in range(LANES):
            hdr, vals = blk[ln]
            if f & F_START:
                cur[ln] = hdr
                acc[ln] = vals[0][0] if init_from_dest else 0.0
                dv[ln] = vals[0][1]
                vals = vals[1:]
            for a, b in vals:
                acc[ln] = step(acc[ln], a, b)
            d, x = cur[ln]
            if lds:
                div = (f & F_END) and (f & F_DIV) and x != 8 * L["ONE"]
                store(sh, d // 8, x // 8 + 1 if div else 0, acc[ln])
            else:
                store(sh, d // 8, x, acc[ln])
        if f & F_END:
            r += 1
    assert r == nr
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
                    if ab is not None and P not in ("fv", "jv"):   # fv, jv read global memory
                        reads += [ab[0], ab[1]]
                for o in reads:
                    if o in last and last[o] >= l:
                        return False
    return True
