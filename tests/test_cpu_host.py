"""CPU-side checks: the C ABI library loads and exports every symbol the
public header declares; mechanism tables are self-consistent; synthetic
inputs have the documented shapes and ranges; the mechanism compiler's
symbolic factorisation reproduces KPP's LU pattern."""
import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parent.parent


def header_symbols():
    txt = (ROOT / "include" / "gckpp_hip.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(gckpp_\w+)\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from gckpp_amd import hip
    if not hip.lib_path().exists():
        pytest.skip("HIP library not built")
    lib = ctypes.CDLL(str(hip.lib_path()))
    syms = header_symbols()
    assert len(syms) >= 14
    for s in syms:
        assert hasattr(lib, s), s
    assert set(syms) <= set(hip.SYMBOLS) | {"gckpp_step_timings"}


def test_fortran_module_record_size_matches_layout():
    """GCKPP_NHIN of the Fortran module is the hetin record of cell_fields.h."""
    from gckpp_amd import fields
    txt = (ROOT / "include" / "gckpp_hip_mod.F90").read_text()
    assert int(re.search(r"GCKPP_NHIN\s*=\s*(\d+)", txt).group(1)) == fields.NHIN


def test_mech_sizes_without_gpu():
    from gckpp_amd import hip, fields
    if not hip.lib_path().exists():
        pytest.skip("HIP library not built")
    assert hip.mech_sizes(0) == [356, 353, 3, 1058, 5683, fields.NMET, fields.NHET, 200]
    assert hip.mech_sizes(1)[:5] == [43, 32, 11, 94, 161]
    assert hip.mech_sizes(2)[:5] == [17, 12, 5, 8, 1]   # KPP/carbon/gckpp_Parameters.F90
    with pytest.raises(hip.GckppError):
        hip.mech_sizes(7)


def test_mechanism_tables_consistent(fullchem):
    m = fullchem
    assert (m.nspec, m.nvar, m.nfix, m.nreact, m.lu_nonzero) == (356, 353, 3, 1058, 5683)
    # every structural Jacobian entry lies in the LU pattern
    pos = set(zip(m.lu_irow.tolist(), m.lu_icol.tolist()))
    nz = set()
    for rx in m.reactions:
        for j, _ in rx["reactants"]:
            if j >= m.nvar:
                continue
            for i, _ in rx["stoich"]:
                nz.add((i, j))
    nz |= {(i, i) for i in range(m.nvar)}     # KPP keeps every diagonal
    assert nz <= pos
    assert len(nz) == m.d["nonzero"]          # KPP's NONZERO (4401)
    # CROW/DIAG describe the row-major pattern with the diagonal present
    for k in range(m.nvar):
        cols = m.lu_icol[m.lu_crow[k]:m.lu_crow[k + 1]]
        assert np.all(np.diff(cols) > 0)
        assert m.lu_icol[m.lu_diag[k]] == k


def test_symbolic_fill_reproduces_kpp_pattern(fullchem):
    """Row-wise symbolic elimination of the Jacobian pattern (with the
    diagonal) in KPP's species order gives exactly KPP's LU pattern."""
    m = fullchem
    rows = [set() for _ in range(m.nvar)]
    for rx in m.reactions:
        for j, _ in rx["reactants"]:
            if j < m.nvar:
                for i, _ in rx["stoich"]:
                    rows[i].add(j)
    for i in range(m.nvar):
        rows[i].add(i)
    filled = []
    for k in range(m.nvar):
        r = set(rows[k])
        for j in sorted(c for c in r if c < k):
            pass
        # eliminate in increasing pivot order, adding U-row patterns
        todo = sorted(c for c in r if c < k)
        done = set()
        while todo:
            j = todo.pop(0)
            if j in done:
                continue
            done.add(j)
            for c in filled[j]:
                if c > j and c not in r:
                    r.add(c)
                    if c < k:
                        todo.append(c)
                        todo.sort()
        filled.append(r)
    pat = set()
    for k in range(m.nvar):
        for c in filled[k]:
            pat.add((k, c))
    assert pat == set(zip(m.lu_irow.tolist(), m.lu_icol.tolist()))


def test_synthetic_box_workload(fullchem):
    from gckpp_amd import synth, fields
    w = synth.box_cells(fullchem, 512, 0)
    T = w.met[fields.MET["TEMP"][0]]
    P = w.met[fields.MET["PRESS"][0]]
    assert T.min() >= 190 and T.max() <= 310
    assert P.min() >= 1.0 and P.max() <= 1013.0
    assert (w.photol[:, :256] == 0).all() and (w.photol[:, 256:] > 0).any()
    assert np.isfinite(w.conc).all() and (w.conc >= 0).all()
    assert w.conc.shape == (356, 512)


def test_grid_workload_shape(fullchem):
    from gckpp_amd import synth, fields
    w = synth.grid_cells(fullchem, 8, 6, 72, seed=1)
    assert w.ncell == 8 * 6 * 72
    P = w.met[fields.MET["PRESS"][0]].reshape(72, 6, 8)
    assert (np.diff(P[:, 0, 0]) < 0).all()         # pressure falls with level
    lit = w.met[fields.MET["SUNCOS"][0]] > 0
    assert 0.2 < lit.mean() < 0.8


def test_fortran_binding_compiles_and_links(tmp_path):
    """include/gckpp_hip_mod.F90 (the ISO_C_BINDING module of INTEGRATION)
    compiles with flang, links against the library and calls
    gckpp_mech_sizes, which needs no GPU."""
    import shutil
    import subprocess
    from gckpp_amd import hip
    flang = shutil.which("flang") or "/opt/rocm/lib/llvm/bin/flang"
    if not Path(flang).exists() or not hip.lib_path().exists():
        pytest.skip("flang or the HIP library not available")
    root = Path(__file__).resolve().parent.parent
    prog = tmp_path / "t.F90"
    prog.write_text(
        "program t\n  use gckpp_hip_mod\n  use iso_c_binding\n  integer(c_int32_t) :: s(8)\n"
        "  integer(c_int) :: rc\n  rc = gckpp_mech_sizes(GCKPP_FULLCHEM, s)\n"
        "  print '(i0,8(1x,i0))', rc, s\nend program t\n")
    libdir = hip.lib_path().parent
    exe = tmp_path / "t"
    subprocess.run([flang, "-o", str(exe), str(root / "include" / "gckpp_hip_mod.F90"), str(prog),
                    "-J", str(tmp_path), "-L" + str(libdir), "-lgckpp_hip",
                    "-Wl,-rpath," + str(libdir)], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split()
    assert [int(v) for v in out[:6]] == [0, 356, 353, 3, 1058, 5683]


def test_flat_segment_tables_reproduce_lu_and_solve(fullchem):
    """The flat segment tables of the persistent solver (Crout U part,
    batched forward/backward substitution) give bit-identical results to the
    position-by-position Crout factorisation and the row-by-row KppSolve
    order on a random matrix with the LU pattern; batches hold only rows that
    do not read each other."""
    import sys
    sys.path.insert(0, str(Path(__file__).resolve().parent.parent / "geos-chem_amd" / "mechgen"))
    from emit_tables import build_tables
    m = fullchem
    raw = {"nvar": m.nvar, "nfix": m.nfix, "nreact": m.nreact, "lu_nonzero": m.lu_nonzero,
           "reactions": m.reactions, "lu_irow": list(m.lu_irow), "lu_icol": list(m.lu_icol),
           "lu_crow": list(m.lu_crow), "lu_diag": list(m.lu_diag)}
    t = build_tables(raw)
    nv, crow, diag, icol = m.nvar, m.lu_crow, m.lu_diag, m.lu_icol
    rng = np.random.default_rng(0)
    A = rng.uniform(-1, 1, m.lu_nonzero)
    A[diag[:nv]] += 40.0
    # reference: Crout position by position
    ref = A.copy()
    for k in range(nv):
        for q in range(crow[k], crow[k + 1]):
            acc = ref[q]
            for tt in range(t["cr_ptr"][q], t["cr_ptr"][q + 1]):
                acc = acc - ref[t["cr_l"][tt]] * ref[t["cr_u"][tt]]
            if t["cr_div"][q] >= 0:
                acc = acc / ref[t["cr_div"][q]]
            ref[q] = acc
    # flat: L part per position, U part by segments
    lu = A.copy()
    for k in range(nv):
        for q in range(crow[k], diag[k]):
            acc = lu[q]
            for tt in range(t["cr_ptr"][q], t["cr_ptr"][q + 1]):
                acc = acc - lu[t["cr_l"][tt]] * lu[t["cr_u"][tt]]
            lu[q] = acc / lu[t["cr_div"][q]]
        out = {}
        for sg in range(t["cu_row"][k], t["cu_row"][k + 1]):
            acc = 0.0
            for tt in range(t["cu_sp"][sg], t["cu_sp"][sg + 1]):
                w = 1.0 if t["cu_l"][tt] == m.lu_nonzero else -lu[t["cu_l"][tt]]
                acc = acc + w * lu[t["cu_u"][tt]]
            out[t["cu_q"][sg]] = acc
        for q, v in out.items():
            lu[q] = v
    assert np.array_equal(lu, ref)
    # solves
    b = rng.normal(size=nv)
    x = b.copy()
    for i in range(nv):
        acc = x[i]
        for k in range(crow[i], diag[i]):
            acc = acc - lu[k] * x[icol[k]]
        x[i] = acc
    for i in range(nv - 1, -1, -1):
        acc = x[i]
        for k in range(diag[i] + 1, crow[i + 1]):
            acc = acc - lu[k] * x[icol[k]]
        x[i] = acc / lu[diag[i]]
    y = b.copy()
    for sp, tl, tx, batch, back in ((t["sf_sp"], t["sf_l"], t["sf_x"], t["sf_batch"], False),
                                    (t["sb_sp"], t["sb_l"], t["sb_x"], t["sb_batch"], True)):
        for bb in range(len(batch) - 1):
            res = {}
            for sg in range(batch[bb], batch[bb + 1]):
                acc, row = 0.0, None
                for tt in range(sp[sg], sp[sg + 1]):
                    if tl[tt] == m.lu_nonzero:
                        row = tx[tt]
                        acc = acc + 1.0 * y[tx[tt]]
                    else:
                        acc = acc - lu[tl[tt]] * y[tx[tt]]
                        # a batch never reads a row it writes
                        assert tx[tt] not in res
                res[row] = acc
            for row, v in res.items():
                y[row] = v / lu[diag[row]] if back else v
    assert np.array_equal(y, x)
    assert len(t["sf_batch"]) - 1 < nv // 4 and len(t["sb_batch"]) - 1 < nv // 4


def test_persistent_variant_count_matches_library():
    from gckpp_amd import hip
    if not hip.lib_path().exists():
        pytest.skip("HIP library not built")
    assert hip.load().gckpp_persistent_variants() == hip.PERSIST_VARIANTS


def test_binding_refuses_host_strided_and_misshaped_tensors():
    """The binding checks every array it hands to the kernels: a host tensor
    or a strided view would be read out of bounds on the device, a wrong
    shape or dtype misread; each is refused with a message before any HIP
    call."""
    import torch
    from gckpp_amd import hip
    with pytest.raises(hip.GckppError, match="device tensor"):
        hip._ptr(torch.zeros(4, dtype=torch.float64))
    assert hip._ptr(None) is None
    t = torch.zeros((3, 8), dtype=torch.float64)
    hip._want(t, "conc", (3, 8), torch.float64)
    hip._want(t, "het", (3, None), torch.float64)
    with pytest.raises(hip.GckppError, match="conc: expected shape"):
        hip._want(t, "conc", (3, 9), torch.float64)
    with pytest.raises(hip.GckppError, match="expected torch.int32"):
        hip._want(t, "ierr", (3, 8), torch.int32)
    with pytest.raises(hip.GckppError, match="expected shape"):
        hip._want(t[:, ::2], "x", (3, 8), torch.float64)


def test_binding_refuses_mismatched_cell_order_and_step_log():
    """A stored cell order or step log indexes the batch it was made for: the
    binding refuses a wrong dtype or rank when it is set, and a persistent
    solve of another cell count afterwards, before any HIP call (checked on a
    context shell that holds no device state)."""
    import torch
    from gckpp_amd import hip
    ctx = object.__new__(hip.Context)
    ctx.torch = torch
    with pytest.raises(hip.GckppError, match="order: expected torch.int32"):
        ctx.set_cell_order(torch.zeros(8, dtype=torch.int64))
    with pytest.raises(hip.GckppError, match="order: expected shape"):
        ctx.set_cell_order(torch.zeros((2, 4), dtype=torch.int32))
    with pytest.raises(hip.GckppError, match="hlog: expected shape"):
        ctx.set_step_log(torch.zeros((6, 8), dtype=torch.float32), 4)
    with pytest.raises(hip.GckppError, match="hlog: expected torch.float32"):
        ctx.set_step_log(torch.zeros((8, 8), dtype=torch.float64), 4)
    with pytest.raises(hip.GckppError, match="nattempt"):
        ctx.set_step_log(torch.zeros((8, 8), dtype=torch.float32), 0)
    ctx._order = torch.zeros(8, dtype=torch.int32)
    ctx._hlog = None
    ctx._check_extras(8, "integrate")
    with pytest.raises(hip.GckppError, match="cell order has 8 entries, the batch 9"):
        ctx._check_extras(9, "integrate")
    ctx._order = None
    ctx._hlog = torch.zeros((8, 8), dtype=torch.float32)
    with pytest.raises(hip.GckppError, match="step log has room for 8 cells"):
        ctx._check_extras(7, "fullchem_step")
