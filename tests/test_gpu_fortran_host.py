"""The library driven from Fortran: geos-chem_amd/fortran/fullchem_host (built
with flang from include/gckpp_hip_mod.F90, the ISO_C_BINDING module, and an
example host program) reads a gathered batch -- C(ncell,NSPEC), MET, PHOTOL,
HET, HSTART as Do_FullChem's gather loop would fill them -- runs one
chemistry step through the host-pointer gckpp_fullchem_step with a context
smaller than the batch (so the step goes in chunks), and writes the
results.  They must equal, bit for bit, the device-pointer step of the
same batch through the Python binding; so must the Python binding's
host-pointer call."""
import subprocess
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = Path(__file__).resolve().parent.parent
HOST = ROOT / "geos-chem_amd" / "fortran" / "fullchem_host"


@pytest.fixture(scope="module")
def torch_dev():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _write_batch(path, m, w, het, max_cells, solver, dt, ic, rc, handoff=0):
    sc = m.species_consts
    with open(path, "wb") as f:
        f.write(np.array([w.ncell, max_cells, het.shape[1]], dtype=np.int64).tobytes())
        f.write(np.array([solver, handoff], dtype=np.int32).tobytes())
        f.write(np.array([dt], dtype=np.float64).tobytes())
        f.write(np.asarray(ic, dtype=np.int32).tobytes())
        f.write(np.asarray(rc, dtype=np.float64).tobytes())
        for a in (m.atol(), m.rtol(), sc["mw"], sc["henry_k0"], sc["henry_cr"]):
            f.write(np.asarray(a, dtype=np.float64).tobytes())
        # [field][cell] in C order is the Fortran A(ncell, NFIELD)
        for a in (w.met, w.photol, het, w.hstart, w.conc):
            f.write(np.ascontiguousarray(a, dtype=np.float64).tobytes())


def _read_result(path, m, n):
    b = open(path, "rb").read()
    off = 0

    def take(dtype, count):
        nonlocal off
        a = np.frombuffer(b, dtype=dtype, count=count, offset=off)
        off += a.nbytes
        return a
    rcode = int(take(np.int32, 1)[0])
    conc = take(np.float64, m.nspec * n).reshape(m.nspec, n)
    hnew = take(np.float64, n)
    ist = take(np.int32, 20 * n).reshape(20, n)
    ierr = take(np.int32, n)
    assert off == len(b)
    return rcode, conc, hnew, ist, ierr


@pytest.mark.parametrize("solver_name,het_per_cell,handoff", [("plain", False, 0),
                                                                ("persistent", True, 0),
                                                                ("persistent", False, 100)])
def test_fortran_host_step_equals_device_step(solver_name, het_per_cell, handoff, fullchem,
                                              torch_dev, tmp_path):
    """(handoff: the host-pointer calls with the tail hand-off of the last
    100 cells of each chunk's persistent launch.)"""
    print("case: Fortran host step, %s solver, per-cell HetState %s, hand-off %d, 600 cells in "
          "chunks of 256" % (solver_name, het_per_cell, handoff), flush=True)
    from gckpp_amd import hip, synth
    from gckpp_amd.options import fullchem_icntrl
    T = torch_dev
    assert HOST.exists(), "the Fortran host is built with the library (make -C geos-chem_amd/csrc)"
    m = fullchem
    solver = {"plain": hip.SOLVER_PLAIN, "persistent": hip.SOLVER_PERSISTENT}[solver_name]
    w = synth.box_cells(m, 600, 23)
    het = np.repeat(w.het, w.ncell, axis=1) if het_per_cell else np.ascontiguousarray(w.het)
    ic, rc = fullchem_icntrl()
    dt = 1200.0

    # the device-pointer step through the Python binding
    ctx = hip.Context(m, 1024)
    ctx.set_solver(solver)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    conc = dev(w.conc)
    hnew, ist, ierr, ok = ctx.fullchem_step(dev(w.met), dev(w.photol), dev(het), conc,
                                            dev(w.hstart), dt, ic, rc)
    want = (conc.cpu().numpy(), hnew.cpu().numpy(), ist.cpu().numpy(), ierr.cpu().numpy())
    assert ok and (want[3] == 1).all()
    # the host-pointer step through the Python binding, chunked
    ctx2 = hip.Context(m, 256)
    ctx2.set_solver(solver)
    ctx2.set_persistent_handoff(handoff)
    c2 = np.ascontiguousarray(w.conc.copy())
    h2, i2, e2, ok2 = ctx2.fullchem_step_host(np.ascontiguousarray(w.met),
                                              np.ascontiguousarray(w.photol), het, c2, w.hstart,
                                              dt, ic, rc)
    assert ok2
    for x, y in zip(want, (c2, h2, i2, e2)):
        assert np.array_equal(x, y)
    ctx.close()
    ctx2.close()

    # the Fortran host, in a child process
    fin, fout = tmp_path / "batch.bin", tmp_path / "result.bin"
    _write_batch(fin, m, w, het, 256, solver, dt, ic, rc, handoff)
    r = subprocess.run([str(HOST), str(fin), str(fout)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    rcode, c3, h3, i3, e3 = _read_result(fout, m, w.ncell)
    assert rcode == 0
    for x, y in zip(want, (c3, h3, i3, e3)):
        assert np.array_equal(x, y)
