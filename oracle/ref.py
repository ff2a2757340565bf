"""ctypes binding of oracle/_ref/libgckpp_ref_<mech>.so (the reference's own
KPP core compiled by oracle/Makefile plus oracle/ref_driver.F90).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg use it, as the checker or the CPU baseline -- never as the
thing measured on the GPU.
"""
import ctypes
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def lib_path(mech="fullchem"):
    return _HERE / "_ref" / ("libgckpp_ref_%s.so" % mech)


def available(mech="fullchem"):
    return lib_path(mech).exists()


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


class RefKPP:
    """The compiled reference for one mechanism.  All per-cell arrays use the
    C-ABI layout [field][ncell] (cells fastest)."""

    def __init__(self, mech):
        self.mech = mech
        self.lib = ctypes.CDLL(str(lib_path(mech.name)))
        sz = (ctypes.c_int * 8)()
        self.lib.ref_sizes(sz)
        self.nspec, self.nvar, self.nfix, self.nreact, self.lu_nonzero = list(sz)[:5]
        assert (self.nspec, self.nvar, self.nreact) == (mech.nspec, mech.nvar, mech.nreact)
        i64 = ctypes.c_int64
        L = self.lib
        L.ref_init.argtypes = [_D] * 5
        L.ref_rates.argtypes = [i64, _D, _D, _D, i64, _D, _D]
        L.ref_fun.argtypes = [i64, _D, _D, _D]
        L.ref_jac.argtypes = [i64, _D, _D, _D]
        L.ref_decomp.argtypes = [i64, _D, _I]
        L.ref_solve.argtypes = [i64, _D, _D]
        L.ref_integrate.argtypes = [i64, _D, _D, _D, ctypes.c_double, ctypes.c_double,
                                    _I, _D, _I, _D, _I]
        L.ref_fullchem_step.argtypes = [i64, _D, _D, _D, i64, _D, _D, ctypes.c_double,
                                        _I, _D, _I, _D, _I, ctypes.c_int]
        sc = mech.species_consts
        self._keep = [np.ascontiguousarray(x, dtype=np.float64) for x in
                      (mech.atol(), mech.rtol(), sc["mw"], sc["henry_k0"], sc["henry_cr"])]
        L.ref_init(*[_d(x) for x in self._keep])

    @staticmethod
    def _het(het, ncell):
        het = np.ascontiguousarray(het, dtype=np.float64)
        return het, (0 if het.shape[1] == 1 else ncell)

    def rates(self, w):
        n = w.ncell
        out = np.zeros((self.nreact, n))
        het, hs = self._het(w.het, n)
        self.lib.ref_rates(n, _d(w.met), _d(w.photol), _d(het), hs, _d(w.conc), _d(out))
        # KPP substitutes constant rates it can fold into Fun/Jac as literals
        # and leaves RCONST(r) unset; fill those so RCONST is complete.
        for r, v in self.mech.inlined_rates().items():
            out[r] = v
        return out

    def fun(self, conc, rconst):
        n = conc.shape[1]
        out = np.zeros((self.nvar, n))
        self.lib.ref_fun(n, _d(np.ascontiguousarray(conc)), _d(np.ascontiguousarray(rconst)), _d(out))
        return out

    def jac(self, conc, rconst):
        n = conc.shape[1]
        out = np.zeros((self.lu_nonzero, n))
        self.lib.ref_jac(n, _d(np.ascontiguousarray(conc)), _d(np.ascontiguousarray(rconst)), _d(out))
        return out

    def decomp(self, jvs):
        lu = np.array(jvs, dtype=np.float64, order="C", copy=True)
        ier = np.zeros(lu.shape[1], dtype=np.int32)
        self.lib.ref_decomp(lu.shape[1], _d(lu), _i(ier))
        return lu, ier

    def solve(self, lu, x):
        x = np.array(x, dtype=np.float64, order="C", copy=True)
        self.lib.ref_solve(x.shape[1], _d(np.ascontiguousarray(lu)), _d(x))
        return x

    def integrate(self, conc, rconst, hstart, tin, tout, icntrl, rcntrl):
        n = conc.shape[1]
        c = np.array(conc, dtype=np.float64, order="C", copy=True)
        ist = np.zeros((20, n), dtype=np.int32)
        rst = np.zeros((20, n))
        ierr = np.zeros(n, dtype=np.int32)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        self.lib.ref_integrate(n, _d(c), _d(np.ascontiguousarray(rconst)),
                               _d(np.ascontiguousarray(hstart, dtype=np.float64)),
                               tin, tout, _i(ic), _d(rc), _i(ist), _d(rst), _i(ierr))
        return c, ist, rst, ierr

    def fullchem_step(self, w, icntrl, rcntrl, nthreads=0):
        """Do_FullChem's per-cell work for every cell of workload w."""
        n = w.ncell
        c = np.array(w.conc, dtype=np.float64, order="C", copy=True)
        ist = np.zeros((20, n), dtype=np.int32)
        hnew = np.zeros(n)
        ierr = np.zeros(n, dtype=np.int32)
        het, hs = self._het(w.het, n)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        self.lib.ref_fullchem_step(n, _d(w.met), _d(w.photol), _d(het), hs, _d(c),
                                   _d(np.ascontiguousarray(w.hstart, dtype=np.float64)),
                                   w.dt, _i(ic), _d(rc), _i(ist), _d(hnew), _i(ierr),
                                   int(nthreads))
        return c, ist, hnew, ierr



# ---- pieces of fullchem_SetStateHet's call tree, compiled from the reference
# (ref_driver.F90 wrappers of rateLawUtilFuncs' Cld_Params and SafeDiv and of
# henry_mod's CALC_KH/CALC_HEFF)

def _vec(*xs):
    n = max(np.size(x) for x in xs)
    return n, [np.ascontiguousarray(np.broadcast_to(np.asarray(x, dtype=np.float64), (n,)))
               for x in xs]


def _flib():
    L = ctypes.CDLL(str(lib_path("fullchem")))
    i64 = ctypes.c_int64
    L.ref_cld_params.argtypes = [i64] + [_D] * 9
    L.ref_safediv.argtypes = [i64] + [_D] * 4
    L.ref_l2g.argtypes = [i64] + [_D] * 7
    L.ref_uptake.argtypes = [i64] + [_D] * 9
    return L


def cld_params(ad, cldf, frland, frocean, qi, ql, t, vair):
    """[6][n]: rLiq, rIce, ALiq, VLiq, AIce, VIce per element."""
    n, a = _vec(ad, cldf, frland, frocean, qi, ql, t, vair)
    out = np.zeros((6, n))
    _flib().ref_cld_params(n, *[_d(x) for x in a], _d(out))
    return out


def safediv(num, den, alt):
    n, a = _vec(num, den, alt)
    out = np.zeros(n)
    _flib().ref_safediv(n, *[_d(x) for x in a], _d(out))
    return out


def l2g(k0, cr, pka, tk, h2oliq, ph):
    n, a = _vec(k0, cr, pka, tk, h2oliq, ph)
    out = np.zeros(n)
    _flib().ref_l2g(n, *[_d(x) for x in a], _d(out))
    return out


def uptake(area, radius, gamma, srmw, numden, temp, cgas, ceduct):
    """[2][n]: Ars_L1k(area, radius, gamma, srMw) and kIIR1Ltd(cgas, ceduct,
    that rate) with NUMDEN, SR_TEMP of each element."""
    n, a = _vec(area, radius, gamma, srmw, numden, temp, cgas, ceduct)
    out = np.zeros((2, n))
    _flib().ref_uptake(n, *[_d(x) for x in a], _d(out))
    return out
