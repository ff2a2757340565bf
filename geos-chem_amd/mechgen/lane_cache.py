"""Mechanism compiler, operand-cache schedules of the persistent lane-per-cell
solver's LU and substitution phases (csrc/rosenbrock.hip, cached variants).

In the lane-per-cell solver every operand of the table-driven sums is one
512-byte access of the wave's slot block in HBM, and the resident slots are
far larger than the L2, so each access is streamed from HBM.  The access
sequence of a phase is fixed by the mechanism's tables and identical for
every lane, so a per-lane cache of S values in LDS can be managed by a
schedule computed here, offline, with Belady's rule (on a miss keep the
value if it is needed again before the cached value needed last): each
operand read carries a descriptor byte saying where it comes from, and each
result a byte saying whether it is also kept.

Descriptor byte: bits 0-4 slot, bits 5-6 kind
  read:   0 MEM   load from the slot block
          1 HIT   read LDS slot
          2 FILL  load from the slot block, then keep it in LDS slot
          3 NONE  (extra operand absent)
  result: 0       written to the slot block only
          2       written to the slot block and kept in LDS slot

Batches (the kernel's execution unit, modelled exactly here): a gather of
n terms runs in batches of 16 while 16 remain, then of 8, 4, 2 and 1 as the
rest allows (`batches`), each of a size fixed at compile time.  All reads
of a batch are issued before any of its fills, so a read sees the cache as
it was when the batch began; the fills follow in read order (never two
fills of one slot in a batch); a sum's result is written after its last
batch.

Records (the descriptor stream, 32-bit words, each record starting on a
word boundary):
  LU section (dev_jac_decomp's positions in order, row by row):
    per row k: Ghimj gather of the diagonal position (kd), then per position
    q of the row: its Ghimj gather (q != kd), then its Crout sum.
    Ghimj gather: per batch of r terms a record of r bytes (one per term's
    partial-derivative operand), padded to a word.
    Crout sum: a first record [div, result, NONE, NONE] + 2r bytes (L
    operand, U operand per term of the first batch), then one record of 2r
    bytes per further batch; n = 0 gives the first record alone.
  substitution section (one KppSolve; the same schedule for every right-hand
  side): forward rows i with a lower part: [x(i), result, NONE, NONE] + 2r
  bytes (LU entry, x(col)) ...; backward rows i = NVAR-1..0: [x(i), result,
  U(i,i), NONE] + 2r bytes ...
Each section starts with an empty cache (a slot is always filled or written
before the schedule reads it), so a section can be restarted at any time
(singular-matrix retries).  Every result also goes to the slot block (LU
entries and x values are read after the section).
"""
SLOTS = (16, 32)   # cache sizes the kernels are built for
MEM, HIT, FILL, NONE = 0, 1, 2, 3
INF = 1 << 60


def batches(n):
    """Batch sizes of an n-term gather, in order."""
    out = [16] * (n // 16)
    for b in (8, 4, 2, 1):
        if n % 16 & b:
            out.append(b)
    return out


def _split(ts):
    out, i = [], 0
    for b in batches(len(ts)):
        out.append(ts[i:i + b])
        i += b
    return out


def lu_events(m, t):
    """The LU section as events: ("G", [addr]) a Ghimj gather batch,
    ("C", first, [addr], extras) a Crout batch (extras: [div addr or None] on
    the first batch of a position), ("W", addr) a result.  Addresses: LU
    entries 0..LU_NZ-1, partial derivatives ("b", p)."""
    ev = []
    nv = m["nvar"]
    crow, diag = m["lu_crow"], m["lu_diag"]

    def ghimj(q):
        for part in _split(list(range(t["jv_ptr"][q], t["jv_ptr"][q + 1]))):
            ev.append(("G", [("b", t["jv_b"][x]) for x in part], []))

    for k in range(nv):
        kd = diag[k]
        ghimj(kd)
        for q in range(crow[k], crow[k + 1]):
            if q != kd:
                ghimj(q)
            dv = t["cr_div"][q]
            parts = _split(list(range(t["cr_ptr"][q], t["cr_ptr"][q + 1]))) or [[]]
            for n_, part in enumerate(parts):
                rd = []
                for x in part:
                    rd += [t["cr_l"][x], t["cr_u"][x]]
                ev.append(("C", rd, [dv if dv >= 0 else None] if n_ == 0 else []))
            ev.append(("W", q))
    return ev


def solve_events(m):
    """One KppSolve as events; x entries are ("x", i), LU entries ints.
    ("S", [addr], extras) a substitution batch (extras on a row's first batch:
    forward [x(i)], backward [x(i), U(i,i)]), ("W", ("x", i)) a result."""
    ev = []
    nv = m["nvar"]
    crow, diag, icol = m["lu_crow"], m["lu_diag"], m["lu_icol"]

    def row(i, lo, hi, extras):
        for n_, part in enumerate(_split(list(range(lo, hi))) or [[]]):
            rd = []
            for k in part:
                rd += [k, ("x", icol[k])]
            ev.append(("S", rd, extras if n_ == 0 else []))
        ev.append(("W", ("x", i)))

    for i in range(nv):
        if crow[i] != diag[i]:
            row(i, crow[i], diag[i], [("x", i)])
    for i in range(nv - 1, -1, -1):
        row(i, diag[i] + 1, crow[i + 1], [("x", i), diag[i]])
    return ev


def plan(ev, S):
    """Belady schedule of an event list with S slots.  Returns per event the
    descriptor bytes: for a batch (extras..., reads...) in the order the
    event lists them (extras first); for a result one byte."""
    # flat access list: (event, addr, is_write)
    acc = []
    for e, x in enumerate(ev):
        if x[0] == "W":
            acc.append((e, x[1], True))
        else:
            for a in x[2]:
                if a is not None:
                    acc.append((e, a, False))
            for a in x[1]:
                acc.append((e, a, False))
    # next read of the value an access sees or leaves (INF: next access is a
    # write, or none)
    nread = [INF] * len(acc)
    nxt = {}
    for i in range(len(acc) - 1, -1, -1):
        a = acc[i][1]
        j = nxt.get(a)
        nread[i] = j if (j is not None and not acc[j][2]) else INF
        nxt[a] = i
    slot_of = {}          # addr -> slot
    nu = [INF] * S        # next read of each slot's value (INF: free/dead)
    owner = [None] * S
    out = []
    i = 0

    def free_slot(exclude):
        best, far = None, -1
        for s in range(S):
            if s in exclude:
                continue
            if owner[s] is None:
                return s, INF + 1
            if nu[s] > far:
                best, far = s, nu[s]
        return best, far

    def take(s, a, n):
        if owner[s] is not None:
            del slot_of[owner[s]]
        owner[s] = a
        slot_of[a] = s
        nu[s] = n

    for e in range(len(ev)):
        j = i
        while j < len(acc) and acc[j][0] == e:
            j += 1
        kind = ev[e][0]
        if kind == "W":
            a = acc[i][1]
            n = nread[i]
            s = slot_of.pop(a, None)
            if s is not None:
                owner[s] = None
                nu[s] = INF
            code = MEM << 5
            if n != INF:
                s2, far = free_slot(())
                if s2 is not None and far > n:
                    take(s2, a, n)
                    code = (FILL << 5) | s2
            out.append([code])
        else:
            codes = []
            miss = []
            for p in range(i, j):
                a = acc[p][1]
                s = slot_of.get(a)
                if s is not None:
                    codes.append((HIT << 5) | s)
                    nu[s] = nread[p]
                else:
                    codes.append(MEM << 5)
                    miss.append(p)
            filled = set()
            for p in miss:
                a = acc[p][1]
                n = nread[p]
                if n == INF or a in slot_of:
                    continue
                s, far = free_slot(filled)
                if s is None or far <= n:
                    continue
                take(s, a, n)
                filled.add(s)
                codes[p - i] = (FILL << 5) | s
            # a slot's next use is that of the batch's last read of its value
            for p in range(i, j):
                s = slot_of.get(acc[p][1])
                if s is not None:
                    nu[s] = nread[p]
            # values read for the last time free their slots
            for s in range(S):
                if owner[s] is not None and nu[s] == INF:
                    del slot_of[owner[s]]
                    owner[s] = None
            out.append(codes)
        i = j
    return out


def encode(ev, codes):
    """Word stream of a section (records as described above)."""
    recs = []
    first = None
    for e, x in enumerate(ev):
        c = codes[e]
        if x[0] == "W":
            recs[first][1] = c[0]        # the result byte of the sum's first record
            continue
        ex = x[2]
        if x[0] == "G" or not ex:
            recs.append(list(c))
            continue
        nx = sum(1 for a in ex if a is not None)
        it = iter(c[:nx])
        xs = [next(it) if a is not None else (NONE << 5) for a in ex]
        recs.append([xs[0], 0, xs[1] if len(xs) > 1 else (NONE << 5), NONE << 5] + list(c[nx:]))
        first = len(recs) - 1
    words = []
    for b in recs:
        b = b + [0] * (-len(b) % 4)
        for q in range(0, len(b), 4):
            words.append(b[q] | (b[q + 1] << 8) | (b[q + 2] << 16) | (b[q + 3] << 24))
    return words


def build(m, t, S):
    """(lu words, solve words) for S slots; m the mechanism dict."""
    if not m["lu_irow"]:
        return [0], [0]
    lev = lu_events(m, t)
    sev = solve_events(m)
    return encode(lev, plan(lev, S)), encode(sev, plan(sev, S))


# ---------------------------------------------------------------------------
# Emulation (tests): execute a section's word stream as the kernel does, on
# versioned symbolic memory, and report any read whose cache slot does not
# hold the current value of the address it stands for.
def emulate(ev, words, S):
    """Returns (mem_reads, hits, mem_writes, errors)."""
    byts = []
    for w in words:
        byts += [w & 255, (w >> 8) & 255, (w >> 16) & 255, w >> 24]
    ver = {}
    lds = [None] * S
    pos = 0
    mem_reads = hits = writes = 0
    errors = []
    pend_result = None

    def read(code, a):
        nonlocal mem_reads, hits
        k, s = code >> 5, code & 31
        cur = (a, ver.get(a, 0))
        if k == HIT:
            hits += 1
            if lds[s] != cur:
                errors.append((a, s, lds[s], cur))
            return None
        mem_reads += 1
        return (s, cur) if k == FILL else None

    for x in ev:
        if x[0] == "W":
            a = x[1]
            ver[a] = ver.get(a, 0) + 1
            writes += 1
            code = pend_result
            if code >> 5 == FILL:
                lds[code & 31] = (a, ver[a])
            continue
        ex = x[2]
        n = len(x[1])
        if x[0] == "G" or not ex:
            body = byts[pos:pos + n]
            pos += (n + 3) & ~3
            exc = []
        else:
            hd = byts[pos:pos + 4]
            body = byts[pos + 4:pos + 4 + n]
            pos += (4 + n + 3) & ~3
            pend_result = hd[1]
            exc = [hd[0]] + ([hd[2]] if len(ex) > 1 else [])
            if len(ex) == 1 and hd[2] >> 5 != NONE:
                errors.append(("extra", hd))
        fills = []
        for c, a in zip(exc, ex if exc else []):
            if a is None:
                if c >> 5 != NONE:
                    errors.append(("extra", a, c))
                continue
            fills.append(read(c, a))
        fills += [read(c, a) for c, a in zip(body, x[1])]
        used = set()
        for f in fills:
            if f is not None:
                s, cur = f
                if s in used:
                    errors.append(("double fill", s))
                used.add(s)
                lds[s] = cur
    return mem_reads, hits, writes, errors
