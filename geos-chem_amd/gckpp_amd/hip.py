"""Host binding of libgckpp_hip.so (include/gckpp_hip.h) for PyTorch device
tensors.  The compute path is the HIP library; there is no CPU fallback: if
the library is missing, loading raises.

Mirrors the reference's operator interface for the path
(KPP/fullchem/gckpp_Rates.F90:408 Update_RCONST and
KPP/fullchem/gckpp_Integrator.F90:80 INTEGRATE), batched over cells:
arrays are [field][ncell] tensors on the GPU (cells fastest).
"""
import ctypes
from pathlib import Path

import numpy as np

_LIB_PATH = Path(__file__).resolve().parent.parent / "csrc" / "libgckpp_hip.so"
_lib = None

SYMBOLS = [
    "gckpp_mech_sizes", "gckpp_init", "gckpp_finalize", "gckpp_set_tolerances",
    "gckpp_set_species_consts", "gckpp_rates_dev", "gckpp_integrate_dev",
    "gckpp_integrate", "gckpp_fullchem_step_dev", "gckpp_fun_dev", "gckpp_jac_dev",
    "gckpp_decomp_dev", "gckpp_solve_dev", "gckpp_step_timings", "gckpp_set_solver",
    "gckpp_last_error",
]
SOLVER_PLAIN, SOLVER_PERSISTENT = 0, 1


class GckppError(RuntimeError):
    pass


def lib_path():
    return _LIB_PATH


def load():
    """Load the HIP library (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not _LIB_PATH.exists():
        raise GckppError("libgckpp_hip.so not built (%s); run __graft_entry__.build()" % _LIB_PATH)
    L = ctypes.CDLL(str(_LIB_PATH))
    vp, i64, i32p, dp = ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p
    L.gckpp_last_error.restype = ctypes.c_char_p
    L.gckpp_mech_sizes.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int32)]
    L.gckpp_init.argtypes = [ctypes.c_int, i64, ctypes.POINTER(vp)]
    L.gckpp_finalize.argtypes = [vp]
    L.gckpp_set_tolerances.argtypes = [vp, dp, dp]
    L.gckpp_set_species_consts.argtypes = [vp, dp, dp, dp]
    L.gckpp_rates_dev.argtypes = [vp, i64, dp, dp, dp, i64, dp, dp, vp]
    L.gckpp_integrate_dev.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp,
                                      dp, dp, dp, i32p, dp, i32p, vp]
    L.gckpp_integrate.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp,
                                  dp, dp, dp, i32p, dp, i32p]
    L.gckpp_fullchem_step_dev.argtypes = [vp, i64, ctypes.c_double, i32p, dp, dp, dp, dp, i64,
                                          dp, dp, dp, i32p, i32p, vp]
    for f in ("gckpp_fun_dev", "gckpp_jac_dev"):
        getattr(L, f).argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_decomp_dev.argtypes = [vp, i64, dp, i32p, vp]
    L.gckpp_solve_dev.argtypes = [vp, i64, dp, dp, vp]
    L.gckpp_step_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.gckpp_set_solver.argtypes = [vp, ctypes.c_int]
    for f in SYMBOLS:
        if f not in ("gckpp_last_error",):
            getattr(L, f).restype = ctypes.c_int
    _lib = L
    return L


def _check(rc, what):
    if rc != 0:
        raise GckppError("%s: %s" % (what, _lib.gckpp_last_error().decode()))


def mech_sizes(mech_id):
    L = load()
    out = (ctypes.c_int32 * 8)()
    _check(L.gckpp_mech_sizes(mech_id, out), "gckpp_mech_sizes")
    return list(out)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def _stream():
    import torch
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class Context:
    """One gckpp context (device workspace for up to max_cells cells)."""

    def __init__(self, mech, max_cells):
        import torch
        self.L = load()
        self.mech = mech
        self.torch = torch
        h = ctypes.c_void_p()
        _check(self.L.gckpp_init(mech.mech_id, int(max_cells), ctypes.byref(h)), "gckpp_init")
        self.h = h
        self.max_cells = int(max_cells)
        self.set_tolerances(mech.atol(), mech.rtol())
        sc = mech.species_consts
        self._keep = [np.ascontiguousarray(sc[k], dtype=np.float64) for k in ("mw", "henry_k0", "henry_cr")]
        _check(self.L.gckpp_set_species_consts(self.h, *[ctypes.c_void_p(a.ctypes.data) for a in self._keep]),
               "gckpp_set_species_consts")

    def close(self):
        if self.h:
            self.L.gckpp_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_solver(self, solver):
        _check(self.L.gckpp_set_solver(self.h, int(solver)), "gckpp_set_solver")

    def set_tolerances(self, atol, rtol):
        a = np.ascontiguousarray(atol, dtype=np.float64)
        r = np.ascontiguousarray(rtol, dtype=np.float64)
        _check(self.L.gckpp_set_tolerances(self.h, ctypes.c_void_p(a.ctypes.data),
                                           ctypes.c_void_p(r.ctypes.data)), "gckpp_set_tolerances")

    # --- Update_RCONST -------------------------------------------------------
    def rates(self, met, photol, het, conc, out=None):
        n = conc.shape[1]
        if out is None:
            out = self.torch.empty((self.mech.nreact, n), dtype=self.torch.float64, device=conc.device)
        hs = 0 if het.shape[1] == 1 else n
        _check(self.L.gckpp_rates_dev(self.h, n, _ptr(met), _ptr(photol), _ptr(het), hs,
                                      _ptr(conc), _ptr(out), _stream()), "gckpp_rates_dev")
        return out

    # --- Integrate -----------------------------------------------------------
    def integrate(self, conc, rconst, hstart, tin, tout, icntrl, rcntrl, istatus=None,
                  rstatus=None, ierr=None):
        """conc is updated in place; returns (istatus, rstatus, ierr) tensors."""
        T = self.torch
        n = conc.shape[1]
        dev = conc.device
        if istatus is None:
            istatus = T.empty((20, n), dtype=T.int32, device=dev)
        if rstatus is None:
            rstatus = T.empty((20, n), dtype=T.float64, device=dev)
        if ierr is None:
            ierr = T.empty((n,), dtype=T.int32, device=dev)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        _check(self.L.gckpp_integrate_dev(self.h, n, float(tin), float(tout),
                                          ctypes.c_void_p(ic.ctypes.data),
                                          ctypes.c_void_p(rc.ctypes.data), _ptr(hstart),
                                          _ptr(conc), _ptr(rconst), _ptr(istatus),
                                          _ptr(rstatus), _ptr(ierr), _stream()),
               "gckpp_integrate_dev")
        return istatus, rstatus, ierr

    def fullchem_step(self, met, photol, het, conc, hstart, dt, icntrl, rcntrl, hnew=None,
                      istatus=None, ierr=None):
        """Do_FullChem per-cell work for all cells; conc updated in place.
        Returns (hnew, istatus, ierr, ok) with ok False if a cell failed twice."""
        T = self.torch
        n = conc.shape[1]
        dev = conc.device
        if hnew is None:
            hnew = T.empty((n,), dtype=T.float64, device=dev)
        if istatus is None:
            istatus = T.empty((20, n), dtype=T.int32, device=dev)
        if ierr is None:
            ierr = T.empty((n,), dtype=T.int32, device=dev)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        hs = 0 if het.shape[1] == 1 else n
        rcode = self.L.gckpp_fullchem_step_dev(self.h, n, float(dt), ctypes.c_void_p(ic.ctypes.data),
                                               ctypes.c_void_p(rc.ctypes.data), _ptr(met), _ptr(photol),
                                               _ptr(het), hs, _ptr(hstart), _ptr(conc), _ptr(hnew),
                                               _ptr(istatus), _ptr(ierr), _stream())
        if rcode != 0:
            msg = self.L.gckpp_last_error().decode()
            if "failed twice" not in msg:
                raise GckppError("gckpp_fullchem_step_dev: " + msg)
        return hnew, istatus, ierr, rcode == 0

    def step_timings(self):
        """(rates_ms, integrate_ms, retry_ms, total_ms) of the last fullchem_step."""
        ms = (ctypes.c_float * 4)()
        _check(self.L.gckpp_step_timings(self.h, ms), "gckpp_step_timings")
        return tuple(float(x) for x in ms)

    # --- components ----------------------------------------------------------
    def fun(self, conc, rconst):
        out = self.torch.empty((self.mech.nvar, conc.shape[1]), dtype=self.torch.float64, device=conc.device)
        _check(self.L.gckpp_fun_dev(self.h, conc.shape[1], _ptr(conc), _ptr(rconst), _ptr(out), _stream()),
               "gckpp_fun_dev")
        return out

    def jac(self, conc, rconst):
        out = self.torch.empty((self.mech.lu_nonzero, conc.shape[1]), dtype=self.torch.float64,
                               device=conc.device)
        _check(self.L.gckpp_jac_dev(self.h, conc.shape[1], _ptr(conc), _ptr(rconst), _ptr(out), _stream()),
               "gckpp_jac_dev")
        return out

    def decomp(self, jvs):
        lu = jvs.clone()
        ier = self.torch.empty((jvs.shape[1],), dtype=self.torch.int32, device=jvs.device)
        _check(self.L.gckpp_decomp_dev(self.h, jvs.shape[1], _ptr(lu), _ptr(ier), _stream()),
               "gckpp_decomp_dev")
        return lu, ier

    def solve(self, lu, x):
        out = x.clone()
        _check(self.L.gckpp_solve_dev(self.h, x.shape[1], _ptr(lu), _ptr(out), _stream()),
               "gckpp_solve_dev")
        return out
