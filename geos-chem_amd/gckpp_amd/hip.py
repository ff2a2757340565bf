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
    "gckpp_set_species_consts", "gckpp_rates_dev", "gckpp_rates", "gckpp_integrate_dev",
    "gckpp_integrate", "gckpp_fullchem_step_dev", "gckpp_fullchem_step", "gckpp_fun_dev",
    "gckpp_jac_dev",
    "gckpp_decomp_dev", "gckpp_solve_dev", "gckpp_step_timings", "gckpp_step_failures",
    "gckpp_set_solver",
    "gckpp_set_persistent_config", "gckpp_get_persistent_config", "gckpp_autotune_dev",
    "gckpp_persistent_variants",
    "gckpp_batch_init", "gckpp_batch_set_config", "gckpp_batch_step_dev",
    "gckpp_batch_capture", "gckpp_batch_replay", "gckpp_batch_finalize",
    "gckpp_set_cell_order", "gckpp_set_step_log",
    "gckpp_mlp_init", "gckpp_mlp_set_norm", "gckpp_mlp_set_params", "gckpp_mlp_get_params",
    "gckpp_mlp_features_dev", "gckpp_mlp_predict_dev", "gckpp_mlp_forward_dev",
    "gckpp_mlp_grad_dev", "gckpp_mlp_train_step_dev", "gckpp_mlp_finalize",
    "gckpp_order_by_class_dev", "gckpp_order_by_h0_dev", "gckpp_order_by_h0_scratch",
    "gckpp_set_persistent_split",
    "gckpp_set_persistent_handoff",
    "gckpp_set_persistent_waves",
    "gckpp_batch_set_handoff", "gckpp_batch_set_waves",
    "gckpp_batch_handoff_count",
    "gckpp_persistent_handoff_stats",
    "gckpp_aout_dev", "gckpp_aout", "gckpp_kppsa_write", "gckpp_kppsa_read",
    "gckpp_kpph_restart_write", "gckpp_kpph_restart_read", "gckpp_integrate_noalloc_dev",
    "gckpp_set_species_map", "gckpp_gather_dev", "gckpp_scatter_dev", "gckpp_set_state_het_dev",
    "gckpp_sulfur_aq_dev", "gckpp_set_state_het", "gckpp_sulfur_aq",
    "gckpp_last_error",
]
MLP_NPARAM = 41360
PERSIST_VARIANTS = 10       # gckpp_persistent_variants(); checked by the CPU tests
WAVE_VARIANT = 7            # one wavefront per cell, state in LDS (64-thread workgroups)


class Part(ctypes.Structure):
    """gckpp_part: the cells of one mechanism in a batch step."""
    _fields_ = [("ncell", ctypes.c_int64), ("met", ctypes.c_void_p), ("photol", ctypes.c_void_p),
                ("het", ctypes.c_void_p), ("het_stride", ctypes.c_int64),
                ("conc", ctypes.c_void_p), ("hstart", ctypes.c_void_p),
                ("istatus", ctypes.c_void_p), ("ierr", ctypes.c_void_p),
                ("icntrl", ctypes.c_int32 * 20), ("rcntrl", ctypes.c_double * 20),
                ("order", ctypes.c_void_p)]
SOLVER_PLAIN, SOLVER_PERSISTENT = 0, 1
NAME_LEN = 64    # GCKPP_NAME_LEN


class KppsaMeta(ctypes.Structure):
    """gckpp_kppsa_meta: the metadata and integrator fields of a
    KPP-Standalone state file."""
    _fields_ = [("location", ctypes.c_char * 256), ("timestamp", ctypes.c_char * 64),
                ("lon", ctypes.c_double), ("lat", ctypes.c_double), ("level", ctypes.c_int32),
                ("pressure_hpa", ctypes.c_double), ("temperature", ctypes.c_double),
                ("airnumden", ctypes.c_double), ("h2o_vmr", ctypes.c_double),
                ("cloud_frac", ctypes.c_double), ("suncos", ctypes.c_double),
                ("h_init", ctypes.c_double), ("h_exit", ctypes.c_double), ("dt", ctypes.c_double),
                ("nsteps", ctypes.c_int32), ("icntrl", ctypes.c_int32 * 20),
                ("rcntrl", ctypes.c_double * 20)]


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
    L.gckpp_rates.argtypes = [vp, i64, dp, dp, dp, i64, dp, dp]
    L.gckpp_fullchem_step.argtypes = [vp, i64, ctypes.c_double, i32p, dp, dp, dp, dp, i64, dp, dp,
                                      dp, i32p, i32p]
    L.gckpp_integrate_dev.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp,
                                      dp, dp, dp, i32p, dp, i32p, vp]
    L.gckpp_integrate_noalloc_dev.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp,
                                              dp, dp, dp, i32p, dp, i32p, vp]
    L.gckpp_integrate.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp,
                                  dp, dp, dp, i32p, dp, i32p]
    L.gckpp_fullchem_step_dev.argtypes = [vp, i64, ctypes.c_double, i32p, dp, dp, dp, dp, i64,
                                          dp, dp, dp, i32p, i32p, vp]
    for f in ("gckpp_fun_dev", "gckpp_jac_dev"):
        getattr(L, f).argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_decomp_dev.argtypes = [vp, i64, dp, i32p, vp]
    L.gckpp_solve_dev.argtypes = [vp, i64, dp, dp, vp]
    L.gckpp_aout_dev.argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_aout.argtypes = [vp, i64, dp, dp, dp]
    L.gckpp_set_species_map.argtypes = [vp, ctypes.c_int32, vp, vp, vp]
    L.gckpp_gather_dev.argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_scatter_dev.argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_sulfur_aq_dev.argtypes = [vp, i64, dp, dp, dp, vp]
    L.gckpp_set_state_het.argtypes = [vp, i64, dp, dp, ctypes.c_int32, ctypes.c_double,
                                      ctypes.c_int32, ctypes.c_int32, dp]
    L.gckpp_sulfur_aq.argtypes = [vp, i64, dp, dp, dp]
    L.gckpp_set_state_het_dev.argtypes = [vp, i64, dp, dp, ctypes.c_int32, ctypes.c_double,
                                          ctypes.c_int32, ctypes.c_int32, dp, vp]
    L.gckpp_kppsa_write.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(KppsaMeta),
                                    ctypes.c_char_p, dp, dp, dp, dp]
    L.gckpp_kpph_restart_write.argtypes = [ctypes.c_char_p, i64, i64, i64, dp, ctypes.c_char_p]
    L.gckpp_kpph_restart_read.argtypes = [ctypes.c_char_p, i64, i64, i64, dp,
                                          ctypes.POINTER(ctypes.c_int32), ctypes.c_char_p]
    L.gckpp_kppsa_read.argtypes = [ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(KppsaMeta),
                                   ctypes.c_char_p, dp, dp, dp, dp]
    L.gckpp_step_timings.argtypes = [vp, ctypes.POINTER(ctypes.c_float)]
    L.gckpp_step_failures.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.gckpp_set_solver.argtypes = [vp, ctypes.c_int]
    L.gckpp_persistent_variants.argtypes = []
    L.gckpp_set_persistent_config.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.gckpp_get_persistent_config.argtypes = [vp, ctypes.POINTER(ctypes.c_int32)]
    L.gckpp_autotune_dev.argtypes = [vp, i64, ctypes.c_double, ctypes.c_double, i32p, dp, dp, dp,
                                     dp, ctypes.POINTER(ctypes.c_float), vp]
    L.gckpp_batch_init.argtypes = [vp, vp, vp, ctypes.POINTER(vp)]
    L.gckpp_batch_set_config.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.gckpp_batch_step_dev.argtypes = [vp, ctypes.POINTER(Part), ctypes.c_double, vp]
    L.gckpp_batch_capture.argtypes = [vp, ctypes.POINTER(Part), ctypes.c_double]
    L.gckpp_batch_replay.argtypes = [vp, vp]
    L.gckpp_batch_finalize.argtypes = [vp]
    L.gckpp_set_cell_order.argtypes = [vp, vp]
    L.gckpp_set_step_log.argtypes = [vp, vp, ctypes.c_int]
    L.gckpp_set_persistent_split.argtypes = [vp, i64]
    L.gckpp_set_persistent_handoff.argtypes = [vp, i64]
    L.gckpp_set_persistent_waves.argtypes = [vp, i64]
    L.gckpp_batch_set_handoff.argtypes = [vp, i64]
    L.gckpp_batch_set_waves.argtypes = [vp, i64]
    L.gckpp_batch_handoff_count.argtypes = [vp, ctypes.POINTER(ctypes.c_int64)]
    L.gckpp_persistent_handoff_stats.argtypes = [vp, ctypes.POINTER(ctypes.c_int64),
                                                 ctypes.POINTER(ctypes.c_float)]
    L.gckpp_order_by_h0_dev.argtypes = [i64, vp, vp, vp, vp]
    L.gckpp_order_by_h0_scratch.argtypes = [i64]
    fp = ctypes.POINTER(ctypes.c_float)
    L.gckpp_mlp_init.argtypes = [i64, vp, vp, ctypes.POINTER(vp)]
    L.gckpp_mlp_set_norm.argtypes = [vp, fp, fp, ctypes.c_float, ctypes.c_float]
    L.gckpp_mlp_set_params.argtypes = [vp, fp]
    L.gckpp_mlp_get_params.argtypes = [vp, fp]
    L.gckpp_mlp_features_dev.argtypes = [vp, i64, dp, dp, dp, dp, vp, vp]
    L.gckpp_mlp_predict_dev.argtypes = [vp, i64, dp, dp, dp, dp, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_double, dp, vp, vp]
    L.gckpp_mlp_forward_dev.argtypes = [vp, i64, vp, vp, vp]
    L.gckpp_mlp_grad_dev.argtypes = [vp, i64, vp, vp, vp, ctypes.c_float, ctypes.c_float, vp, vp, vp]
    L.gckpp_mlp_train_step_dev.argtypes = [vp, i64, vp, vp, vp, ctypes.c_float, ctypes.c_float,
                                           ctypes.c_float, vp, vp]
    L.gckpp_mlp_finalize.argtypes = [vp]
    L.gckpp_order_by_class_dev.argtypes = [i64, vp, vp, vp, vp]
    for f in SYMBOLS:
        if f not in ("gckpp_last_error", "gckpp_order_by_h0_scratch"):
            getattr(L, f).restype = ctypes.c_int
    L.gckpp_order_by_h0_scratch.restype = ctypes.c_int64
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
    """Device address of a tensor (None: a null pointer).  The kernels index
    every array themselves, so a host or strided tensor is refused here
    rather than read out of bounds on the device."""
    if t is None:
        return None
    if not t.is_cuda:
        raise GckppError("expected a device tensor, got one on %s" % t.device)
    if not t.is_contiguous():
        raise GckppError("expected a contiguous tensor, got strides %s" % (t.stride(),))
    return ctypes.c_void_p(t.data_ptr())


def _want(t, name, shape, dtype):
    """Checks a caller's array: shape (None entries: any) and dtype."""
    if t is None:
        return
    ok = t.dim() == len(shape) and all(w is None or w == g for w, g in zip(shape, t.shape))
    if not ok:
        raise GckppError("%s: expected shape %s, got %s" % (name, tuple(shape), tuple(t.shape)))
    if t.dtype != dtype:
        raise GckppError("%s: expected %s, got %s" % (name, dtype, t.dtype))


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
        sz = mech_sizes(mech.mech_id)
        self.nmet, self.nhet, self.nphotol = sz[5], sz[6], sz[7]
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

    def set_persistent_config(self, variant, waves_per_group=1):
        _check(self.L.gckpp_set_persistent_config(self.h, int(variant), int(waves_per_group)),
               "gckpp_set_persistent_config")

    def set_cell_order(self, order):
        """Order (device int32 permutation of 0..n-1, or None) in which the
        persistent solver takes cells; the tensor must outlive its use.  Later
        persistent solves refuse a batch whose cell count differs from its
        length (the kernels would read past its end)."""
        if order is not None:
            _want(order, "order", (None,), self.torch.int32)
        p = _ptr(order)
        _check(self.L.gckpp_set_cell_order(self.h, p), "gckpp_set_cell_order")
        self._order = order

    def set_persistent_split(self, nfirst):
        """The first nfirst cells of the cell order run on the cell-per-wave
        kernel on a second stream, concurrently with the lane kernel on the
        rest (0: off); see gckpp_set_persistent_split."""
        _check(self.L.gckpp_set_persistent_split(self.h, int(nfirst)), "gckpp_set_persistent_split")

    def set_persistent_handoff(self, max_cells):
        """Once at most max_cells cells of a lane-kernel launch are
        unfinished, they move at a step boundary to a cell-per-wave launch
        that follows it (0: off); see gckpp_set_persistent_handoff."""
        _check(self.L.gckpp_set_persistent_handoff(self.h, int(max_cells)),
               "gckpp_set_persistent_handoff")

    def set_persistent_waves(self, max_waves):
        """At most max_waves resident waves in a persistent launch (0: all)."""
        _check(self.L.gckpp_set_persistent_waves(self.h, int(max_waves)),
               "gckpp_set_persistent_waves")

    def handoff_stats(self):
        """(cells handed over, attempts the cell-per-wave launch made on them,
        lane launch ms, cell-per-wave launch ms) of the last persistent solve
        with a hand-off."""
        n = (ctypes.c_int64 * 2)()
        ms = (ctypes.c_float * 2)()
        _check(self.L.gckpp_persistent_handoff_stats(self.h, n, ms),
               "gckpp_persistent_handoff_stats")
        return int(n[0]), int(n[1]), float(ms[0]), float(ms[1])

    def set_step_log(self, hlog, nattempt=0):
        """Device float32 tensor [2*nattempt][ncell] (or None) receiving (H,
        err) of each cell's first attempts in the next persistent solves,
        which refuse a batch of another cell count."""
        if hlog is not None:
            if int(nattempt) <= 0:
                raise GckppError("set_step_log: nattempt must be positive with a log")
            _want(hlog, "hlog", (2 * int(nattempt), None), self.torch.float32)
        p = _ptr(hlog)
        _check(self.L.gckpp_set_step_log(self.h, p, int(nattempt)), "gckpp_set_step_log")
        self._hlog = hlog

    def _check_extras(self, n, who):
        """The stored cell order and step log index a batch of their own
        length: refuse a persistent solve of n cells that does not match."""
        o = getattr(self, "_order", None)
        if o is not None and o.numel() != n:
            raise GckppError("%s: the cell order has %d entries, the batch %d cells"
                             % (who, o.numel(), n))
        g = getattr(self, "_hlog", None)
        if g is not None and g.shape[1] != n:
            raise GckppError("%s: the step log has room for %d cells, the batch has %d"
                             % (who, g.shape[1], n))

    def persistent_config(self):
        """dict(variant, waves_per_group, waves, loads_in_flight,
        min_waves_per_simd, segments, wave_per_cell, cache_slots)."""
        out = (ctypes.c_int32 * 7)()
        _check(self.L.gckpp_get_persistent_config(self.h, out), "gckpp_get_persistent_config")
        d = dict(zip(("variant", "waves_per_group", "waves", "loads_in_flight",
                      "min_waves_per_simd"), list(out)[:5]))
        d["segments"] = int(out[5] == 1)
        d["wave_per_cell"] = int(out[5] == 2)
        d["cache_slots"] = int(out[6])
        return d

    def autotune(self, conc, rconst, hstart, tin, tout, icntrl, rcntrl):
        """Times every persistent configuration on these cells (conc is not
        modified), keeps the fastest; returns ms[variant][log2 waves_per_group]."""
        nv = self.L.gckpp_persistent_variants()
        assert nv == PERSIST_VARIANTS, "library and binding disagree on the variant count"
        n = conc.shape[1]
        _want(conc, "conc", (self.mech.nspec, n), self.torch.float64)
        _want(rconst, "rconst", (self.mech.nreact, n), self.torch.float64)
        _want(hstart, "hstart", (n,), self.torch.float64)
        self._check_extras(n, "autotune")
        ms = (ctypes.c_float * (3 * nv))()
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        _check(self.L.gckpp_autotune_dev(self.h, conc.shape[1], float(tin), float(tout),
                                         ctypes.c_void_p(ic.ctypes.data),
                                         ctypes.c_void_p(rc.ctypes.data), _ptr(hstart),
                                         _ptr(conc), _ptr(rconst), ms, _stream()),
               "gckpp_autotune_dev")
        return np.array(list(ms), dtype=np.float32).reshape(PERSIST_VARIANTS, 3)

    def set_tolerances(self, atol, rtol):
        a = np.ascontiguousarray(atol, dtype=np.float64)
        r = np.ascontiguousarray(rtol, dtype=np.float64)
        _check(self.L.gckpp_set_tolerances(self.h, ctypes.c_void_p(a.ctypes.data),
                                           ctypes.c_void_p(r.ctypes.data)), "gckpp_set_tolerances")

    # --- Update_RCONST -------------------------------------------------------
    def _inputs(self, met, photol, het, conc):
        """Checks the per-cell input records; returns ncell."""
        T, m = self.torch, self.mech
        n = conc.shape[1]
        _want(conc, "conc", (m.nspec, n), T.float64)
        _want(met, "met", (self.nmet, n), T.float64)
        _want(photol, "photol", (self.nphotol, n), T.float64)
        _want(het, "het", (self.nhet, None), T.float64)
        if het.shape[1] not in (1, n):
            raise GckppError("het: expected 1 or %d columns, got %d" % (n, het.shape[1]))
        return n

    def rates(self, met, photol, het, conc, out=None):
        n = self._inputs(met, photol, het, conc)
        if out is None:
            out = self.torch.empty((self.mech.nreact, n), dtype=self.torch.float64, device=conc.device)
        _want(out, "rconst", (self.mech.nreact, n), self.torch.float64)
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
        _want(conc, "conc", (self.mech.nspec, n), T.float64)
        _want(rconst, "rconst", (self.mech.nreact, n), T.float64)
        _want(hstart, "hstart", (n,), T.float64)
        _want(istatus, "istatus", (20, n), T.int32)
        _want(rstatus, "rstatus", (20, n), T.float64)
        _want(ierr, "ierr", (n,), T.int32)
        self._check_extras(n, "integrate")
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
        n = self._inputs(met, photol, het, conc)
        dev = conc.device
        _want(hstart, "hstart", (n,), T.float64)
        _want(hnew, "hnew", (n,), T.float64)
        _want(istatus, "istatus", (20, n), T.int32)
        _want(ierr, "ierr", (n,), T.int32)
        self._check_extras(n, "fullchem_step")
        if hnew is None:
            hnew = T.empty((n,), dtype=T.float64, device=dev)
        if istatus is None:
            istatus = T.empty((20, n), dtype=T.int32, device=dev)
        if ierr is None:
            ierr = T.empty((n,), dtype=T.int32, device=dev)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        hs = 0 if het.shape[1] == 1 else n
        _check(self.L.gckpp_fullchem_step_dev(self.h, n, float(dt), ctypes.c_void_p(ic.ctypes.data),
                                              ctypes.c_void_p(rc.ctypes.data), _ptr(met), _ptr(photol),
                                              _ptr(het), hs, _ptr(hstart), _ptr(conc), _ptr(hnew),
                                              _ptr(istatus), _ptr(ierr), _stream()),
               "gckpp_fullchem_step_dev")
        return hnew, istatus, ierr, self.step_failures() == 0

    def step_failures(self):
        """Cells of the last fullchem_step that failed twice (synchronises)."""
        v = ctypes.c_int64(0)
        _check(self.L.gckpp_step_failures(self.h, ctypes.byref(v)), "gckpp_step_failures")
        return int(v.value)

    def fullchem_step_host(self, met, photol, het, conc, hstart, dt, icntrl, rcntrl):
        """gckpp_fullchem_step on host (numpy) arrays, the Fortran host's call:
        conc [NSPEC][n] float64 updated in place; returns (hnew, istatus,
        ierr, ok) as numpy arrays."""
        m = self.mech
        n = conc.shape[1]
        arrs = {"met": (met, (self.nmet, n)), "photol": (photol, (self.nphotol, n)),
                "conc": (conc, (m.nspec, n))}
        for name, (a, shp) in arrs.items():
            if a.dtype != np.float64 or a.shape != shp or not a.flags.c_contiguous:
                raise GckppError("%s: expected a C-contiguous float64 array of shape %s" % (name, shp))
        if het.dtype != np.float64 or het.shape[0] != self.nhet or het.shape[1] not in (1, n) \
                or not het.flags.c_contiguous:
            raise GckppError("het: expected a C-contiguous float64 array [%d][1 or %d]" % (self.nhet, n))
        hs = None
        if hstart is not None:
            hs = np.ascontiguousarray(hstart, dtype=np.float64)
            if hs.shape != (n,):
                raise GckppError("hstart: expected shape (%d,)" % n)
        hnew = np.zeros(n)
        ist = np.zeros((20, n), dtype=np.int32)
        ierr = np.zeros(n, dtype=np.int32)
        ic = np.ascontiguousarray(icntrl, dtype=np.int32)
        rc = np.ascontiguousarray(rcntrl, dtype=np.float64)
        h = lambda a: None if a is None else ctypes.c_void_p(a.ctypes.data)
        rcode = self.L.gckpp_fullchem_step(self.h, n, float(dt), h(ic), h(rc), h(met), h(photol),
                                           h(het), 0 if het.shape[1] == 1 else n, h(hs), h(conc),
                                           h(hnew), h(ist), h(ierr))
        if rcode != 0:
            msg = self.L.gckpp_last_error().decode()
            if "failed twice" not in msg:
                raise GckppError("gckpp_fullchem_step: " + msg)
        return hnew, ist, ierr, rcode == 0

    def step_timings(self):
        """(rates_ms, integrate_ms, retry_ms, total_ms) of the last fullchem_step."""
        ms = (ctypes.c_float * 4)()
        _check(self.L.gckpp_step_timings(self.h, ms), "gckpp_step_timings")
        return tuple(float(x) for x in ms)

    # --- components ----------------------------------------------------------
    def _comp_in(self, conc, rconst):
        n = conc.shape[1]
        _want(conc, "conc", (self.mech.nspec, n), self.torch.float64)
        _want(rconst, "rconst", (self.mech.nreact, n), self.torch.float64)

    def fun(self, conc, rconst):
        self._comp_in(conc, rconst)
        out = self.torch.empty((self.mech.nvar, conc.shape[1]), dtype=self.torch.float64, device=conc.device)
        _check(self.L.gckpp_fun_dev(self.h, conc.shape[1], _ptr(conc), _ptr(rconst), _ptr(out), _stream()),
               "gckpp_fun_dev")
        return out

    def jac(self, conc, rconst):
        self._comp_in(conc, rconst)
        out = self.torch.empty((self.mech.lu_nonzero, conc.shape[1]), dtype=self.torch.float64,
                               device=conc.device)
        _check(self.L.gckpp_jac_dev(self.h, conc.shape[1], _ptr(conc), _ptr(rconst), _ptr(out), _stream()),
               "gckpp_jac_dev")
        return out

    def decomp(self, jvs):
        _want(jvs, "jvs", (self.mech.lu_nonzero, jvs.shape[1]), self.torch.float64)
        lu = jvs.clone()
        ier = self.torch.empty((jvs.shape[1],), dtype=self.torch.int32, device=jvs.device)
        _check(self.L.gckpp_decomp_dev(self.h, jvs.shape[1], _ptr(lu), _ptr(ier), _stream()),
               "gckpp_decomp_dev")
        return lu, ier

    def solve(self, lu, x):
        _want(lu, "lu", (self.mech.lu_nonzero, x.shape[1]), self.torch.float64)
        _want(x, "x", (self.mech.nvar, x.shape[1]), self.torch.float64)
        out = x.clone()
        _check(self.L.gckpp_solve_dev(self.h, x.shape[1], _ptr(lu), _ptr(out), _stream()),
               "gckpp_solve_dev")
        return out

    def set_species_map(self, model_mw_g, kpp_to_model, zero=None, alk_equiv=False):
        """Map_KppSpc: kpp_to_model[NSPEC] (model species index or -1),
        model_mw_g[nmodel] [g/mol], zero[NSPEC] the prod/loss dummies;
        alk_equiv: SALAAL/SALCAL to equivalents in the gather and back in
        the scatter (fullchem)."""
        mp = np.ascontiguousarray(kpp_to_model, dtype=np.int32)
        mw = np.ascontiguousarray(model_mw_g, dtype=np.float64)
        if mp.shape != (self.mech.nspec,):
            raise GckppError("set_species_map: map of shape %s" % (mp.shape,))
        zz = None if zero is None else np.ascontiguousarray(zero, dtype=np.int32)
        _check(self.L.gckpp_set_species_map(self.h, int(mw.size), mp.ctypes.data_as(ctypes.c_void_p),
                                            mw.ctypes.data_as(ctypes.c_void_p),
                                            None if zz is None else zz.ctypes.data_as(ctypes.c_void_p),
                                            int(bool(alk_equiv))),
               "gckpp_set_species_map")

    def gather(self, species, airden, conc=None):
        """species[nmodel][ncell] kg/kg dry, airden[ncell] kg/m3 ->
        conc[NSPEC][ncell] molec/cm3 (device tensors)."""
        T = self.torch
        n = species.shape[1]
        _want(species, "species", (species.shape[0], n), T.float64)
        _want(airden, "airden", (n,), T.float64)
        if conc is None:
            conc = T.empty((self.mech.nspec, n), dtype=T.float64, device=species.device)
        _want(conc, "conc", (self.mech.nspec, n), T.float64)
        _check(self.L.gckpp_gather_dev(self.h, n, _ptr(species), _ptr(airden), _ptr(conc), _stream()),
               "gckpp_gather_dev")
        return conc

    def scatter(self, conc, airden, species):
        """The mapped species of conc back into species (kg/kg dry)."""
        T = self.torch
        n = conc.shape[1]
        _want(conc, "conc", (self.mech.nspec, n), T.float64)
        _want(airden, "airden", (n,), T.float64)
        _want(species, "species", (species.shape[0], n), T.float64)
        _check(self.L.gckpp_scatter_dev(self.h, n, _ptr(conc), _ptr(airden), _ptr(species), _stream()),
               "gckpp_scatter_dev")

    def set_state_het(self, hetin, conc, het=None, naerotype=14, gamma_ho2=0.2, lpscchem=True,
                      turn_off_het=False):
        """fullchem_SetStateHet: het[NHET][ncell] from the model fields
        hetin[NHIN][ncell] and conc[NSPEC][ncell] (device tensors)."""
        from . import fields
        T = self.torch
        n = conc.shape[1]
        _want(hetin, "hetin", (fields.NHIN, n), T.float64)
        _want(conc, "conc", (self.mech.nspec, n), T.float64)
        if het is None:
            het = T.empty((fields.NHET, n), dtype=T.float64, device=conc.device)
        _want(het, "het", (fields.NHET, n), T.float64)
        _check(self.L.gckpp_set_state_het_dev(self.h, n, _ptr(hetin), _ptr(conc), int(naerotype),
                                              float(gamma_ho2), int(bool(lpscchem)),
                                              int(bool(turn_off_het)), _ptr(het), _stream()),
               "gckpp_set_state_het_dev")
        return het

    def sulfur_aq(self, hetin, conc, met):
        """fullchem_SulfurAqChem: K_MT(1:6) into met's rows MET_K_MT.. in
        place (device tensors)."""
        from . import fields
        T = self.torch
        n = conc.shape[1]
        _want(hetin, "hetin", (fields.NHIN, n), T.float64)
        _want(conc, "conc", (self.mech.nspec, n), T.float64)
        _want(met, "met", (fields.NMET, n), T.float64)
        _check(self.L.gckpp_sulfur_aq_dev(self.h, n, _ptr(hetin), _ptr(conc), _ptr(met), _stream()),
               "gckpp_sulfur_aq_dev")
        return met

    def set_state_het_host(self, hetin, conc, naerotype=14, gamma_ho2=0.2, lpscchem=True,
                           turn_off_het=False):
        """The host-pointer fullchem_SetStateHet: numpy hetin[NHIN][n],
        conc[NSPEC][n] -> het[NHET][n]."""
        from . import fields
        hetin = np.ascontiguousarray(hetin, dtype=np.float64)
        conc = np.ascontiguousarray(conc, dtype=np.float64)
        n = conc.shape[1]
        if hetin.shape != (fields.NHIN, n) or conc.shape[0] != self.mech.nspec:
            raise GckppError("set_state_het_host: shapes %s, %s" % (hetin.shape, conc.shape))
        het = np.empty((fields.NHET, n))
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _check(self.L.gckpp_set_state_het(self.h, n, P(hetin), P(conc), int(naerotype),
                                          float(gamma_ho2), int(bool(lpscchem)),
                                          int(bool(turn_off_het)), P(het)), "gckpp_set_state_het")
        return het

    def sulfur_aq_host(self, hetin, conc, met):
        """The host-pointer fullchem_SulfurAqChem: K_MT into a copy of met."""
        from . import fields
        hetin = np.ascontiguousarray(hetin, dtype=np.float64)
        conc = np.ascontiguousarray(conc, dtype=np.float64)
        out = np.array(met, dtype=np.float64, order="C", copy=True)
        n = conc.shape[1]
        if hetin.shape != (fields.NHIN, n) or out.shape != (fields.NMET, n):
            raise GckppError("sulfur_aq_host: shapes %s, %s" % (hetin.shape, out.shape))
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        _check(self.L.gckpp_sulfur_aq(self.h, n, P(hetin), P(conc), P(out)), "gckpp_sulfur_aq")
        return out

    def aout(self, conc, rconst):
        """Fun's reaction rates A[NREACT][ncell]."""
        self._comp_in(conc, rconst)
        out = self.torch.empty((self.mech.nreact, conc.shape[1]), dtype=self.torch.float64,
                               device=conc.device)
        _check(self.L.gckpp_aout_dev(self.h, conc.shape[1], _ptr(conc), _ptr(rconst), _ptr(out),
                                     _stream()), "gckpp_aout_dev")
        return out


class Batch:
    """gckpp_batch: one persistent launch per chemistry step over the cells of
    up to three mechanisms (contexts in the order fullchem, Hg, carbon)."""

    def __init__(self, fullchem=None, hg=None, carbon=None):
        self.L = load()
        self.ctxs = [fullchem, hg, carbon]
        h = ctypes.c_void_p()
        _check(self.L.gckpp_batch_init(*[c.h if c is not None else None for c in self.ctxs],
                                       ctypes.byref(h)), "gckpp_batch_init")
        self.h = h
        self._parts = None

    def set_config(self, variant, waves_per_group=1):
        _check(self.L.gckpp_batch_set_config(self.h, int(variant), int(waves_per_group)),
               "gckpp_batch_set_config")

    def set_waves(self, max_waves):
        """Cap on the persistent launch's waves (0: every resident wave)."""
        _check(self.L.gckpp_batch_set_waves(self.h, int(max_waves)), "gckpp_batch_set_waves")

    @staticmethod
    def part(met=None, photol=None, het=None, conc=None, hstart=None, icntrl=None, rcntrl=None,
             istatus=None, ierr=None, order=None):
        """A gckpp_part over device tensors (conc [NSPEC][n], hstart [n] in/out)."""
        p = Part()
        if conc is None:
            return p
        n = conc.shape[1]
        p.ncell = n
        p.met, p.photol, p.het = _ptr(met), _ptr(photol), _ptr(het)
        p.het_stride = 0 if het.shape[1] == 1 else n
        p.conc, p.hstart = _ptr(conc), _ptr(hstart)
        p.istatus, p.ierr = _ptr(istatus), _ptr(ierr)
        p.order = _ptr(order)
        for k in range(20):
            p.icntrl[k] = int(icntrl[k])
            p.rcntrl[k] = float(rcntrl[k])
        return p

    def _arr(self, parts):
        arr = (Part * 3)(*[parts[k] if k < len(parts) and parts[k] is not None else Part()
                           for k in range(3)])
        self._parts = arr          # the graph keeps the pointers; keep the tensors alive too
        return arr

    def set_handoff(self, max_cells):
        """Tail hand-off of the batch launch (0: off)."""
        _check(self.L.gckpp_batch_set_handoff(self.h, int(max_cells)), "gckpp_batch_set_handoff")

    def handoff_count(self):
        n = ctypes.c_int64(0)
        _check(self.L.gckpp_batch_handoff_count(self.h, ctypes.byref(n)), "gckpp_batch_handoff_count")
        return n.value

    def step(self, parts, dt):
        _check(self.L.gckpp_batch_step_dev(self.h, self._arr(parts), float(dt), _stream()),
               "gckpp_batch_step_dev")

    def capture(self, parts, dt):
        _check(self.L.gckpp_batch_capture(self.h, self._arr(parts), float(dt)),
               "gckpp_batch_capture")

    def replay(self):
        _check(self.L.gckpp_batch_replay(self.h, _stream()), "gckpp_batch_replay")

    def close(self):
        if self.h:
            self.L.gckpp_batch_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def kppsa_write(mech, path, meta, conc, atol, rconst, aout, names=None):
    """Writes one cell's KPP-Standalone state file (host numpy arrays:
    conc[NSPEC], atol[NSPEC], rconst[NREACT], aout[NREACT]; names: NSPEC
    strings, None for "C<n>" lines)."""
    import numpy as np
    L = load()
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (conc, atol, rconst, aout)]
    for a, n in zip(arrs, (mech.nspec, mech.nspec, mech.nreact, mech.nreact)):
        if a.shape != (n,):
            raise GckppError("kppsa_write: array of shape %s, expected (%d,)" % (a.shape, n))
    nb = None
    if names is not None:
        if len(names) != mech.nspec:
            raise GckppError("kppsa_write: %d names for %d species" % (len(names), mech.nspec))
        nb = b"".join((n or "").encode()[:NAME_LEN - 1].ljust(NAME_LEN, b"\0") for n in names)
    _check(L.gckpp_kppsa_write(mech.mech_id, str(path).encode(), ctypes.byref(meta), nb,
                               *[a.ctypes.data_as(ctypes.c_void_p) for a in arrs]), "gckpp_kppsa_write")


def kppsa_read(mech, path):
    """Reads a KPP-Standalone state file: (meta, names, conc, atol, rconst, aout)."""
    import numpy as np
    L = load()
    meta = KppsaMeta()
    nb = ctypes.create_string_buffer(mech.nspec * NAME_LEN)
    conc, atol = np.zeros(mech.nspec), np.zeros(mech.nspec)
    rct, aout = np.zeros(mech.nreact), np.zeros(mech.nreact)
    _check(L.gckpp_kppsa_read(mech.mech_id, str(path).encode(), ctypes.byref(meta), nb,
                              *[a.ctypes.data_as(ctypes.c_void_p) for a in (conc, atol, rct, aout)]),
           "gckpp_kppsa_read")
    raw = nb.raw
    names = [raw[i * NAME_LEN:(i + 1) * NAME_LEN].split(b"\0", 1)[0].decode() for i in range(mech.nspec)]
    return meta, names, conc, atol, rct, aout


def kpph_restart_write(path, kpph, model_tag=""):
    """Writes KPPHvalue (numpy array (nx, ny, nz), any memory order) as the
    restart variable Chem_KPPHvalue."""
    import numpy as np
    L = load()
    k = np.asfortranarray(kpph, dtype=np.float64)
    nx, ny, nz = k.shape
    flat = np.ascontiguousarray(k.reshape(-1, order="F"))
    _check(L.gckpp_kpph_restart_write(str(path).encode(), nx, ny, nz,
                                      flat.ctypes.data_as(ctypes.c_void_p), model_tag.encode()),
           "gckpp_kpph_restart_write")


def kpph_restart_read(path, shape):
    """(KPPHvalue (nx, ny, nz), found, model tag) from a restart file."""
    import numpy as np
    L = load()
    nx, ny, nz = shape
    flat = np.zeros(nx * ny * nz)
    found = ctypes.c_int32()
    tag = ctypes.create_string_buffer(NAME_LEN)
    _check(L.gckpp_kpph_restart_read(str(path).encode(), nx, ny, nz,
                                     flat.ctypes.data_as(ctypes.c_void_p), ctypes.byref(found), tag),
           "gckpp_kpph_restart_read")
    return flat.reshape((nx, ny, nz), order="F"), bool(found.value), tag.value.decode()


def order_by_h0(h0):
    """Cells sorted by predicted h0, smallest (stiffest) first, in quarter-
    decade bins, stable: int32 device tensor."""
    import torch
    n = h0.shape[0]
    if h0.dtype != torch.float64:
        raise GckppError("h0: expected torch.float64, got %s" % h0.dtype)
    L = load()
    order = torch.empty((n,), dtype=torch.int32, device=h0.device)
    scratch = torch.empty((max(1, int(L.gckpp_order_by_h0_scratch(n))),), dtype=torch.int32,
                          device=h0.device)
    _check(L.gckpp_order_by_h0_dev(n, _ptr(h0), _ptr(order), _ptr(scratch), _stream()),
           "gckpp_order_by_h0_dev")
    return order


def order_by_class(cls):
    """Cells sorted by class, costliest (3) first, stable: int32 device tensor."""
    import torch
    n = cls.shape[0]
    order = torch.empty((n,), dtype=torch.int32, device=cls.device)
    scratch = torch.empty((4 * ((n + 255) // 256) + 4,), dtype=torch.int32, device=cls.device)
    _check(load().gckpp_order_by_class_dev(n, _ptr(cls), _ptr(order), _ptr(scratch), _stream()),
           "gckpp_order_by_class_dev")
    return order


class MLP:
    """gckpp_mlp: the step-size / stiffness-class network (bf16 MFMA
    inference, fp32 training kernels)."""

    def __init__(self, max_batch, jslot, spc):
        import torch
        self.L = load()
        self.torch = torch
        j = np.ascontiguousarray(jslot, dtype=np.int32)
        sp = np.ascontiguousarray(spc, dtype=np.int32)
        assert j.size == 12 and sp.size == 24
        h = ctypes.c_void_p()
        _check(self.L.gckpp_mlp_init(int(max_batch), ctypes.c_void_p(j.ctypes.data),
                                     ctypes.c_void_p(sp.ctypes.data), ctypes.byref(h)),
               "gckpp_mlp_init")
        self.h = h
        self.max_batch = int(max_batch)
        self._loss = torch.zeros(1, dtype=torch.float32, device="cuda")

    @staticmethod
    def _f(a):
        a = np.ascontiguousarray(a, dtype=np.float32)
        return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))

    def set_norm(self, mean, std, out_mu, out_sigma):
        m, mp = self._f(mean)
        s, sp = self._f(std)
        assert m.size == 48 and s.size == 48
        _check(self.L.gckpp_mlp_set_norm(self.h, mp, sp, float(out_mu), float(out_sigma)),
               "gckpp_mlp_set_norm")

    def set_params(self, params):
        p, pp = self._f(params)
        assert p.size == MLP_NPARAM
        _check(self.L.gckpp_mlp_set_params(self.h, pp), "gckpp_mlp_set_params")

    def params(self):
        out = np.zeros(MLP_NPARAM, dtype=np.float32)
        _check(self.L.gckpp_mlp_get_params(self.h, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float))),
               "gckpp_mlp_get_params")
        return out

    def features(self, met, photol, conc, hprev=None):
        n = conc.shape[1]
        X = self.torch.empty((n, 48), dtype=self.torch.float32, device=conc.device)
        _check(self.L.gckpp_mlp_features_dev(self.h, n, _ptr(met), _ptr(photol), _ptr(conc),
                                             _ptr(hprev), _ptr(X), _stream()),
               "gckpp_mlp_features_dev")
        return X

    def predict(self, met, photol, conc, hprev=None, safety=1.0, hmin=1e-5, hmax=1200.0, h0=None,
                cls=None):
        T = self.torch
        n = conc.shape[1]
        if h0 is None:
            h0 = T.empty((n,), dtype=T.float64, device=conc.device)
        if cls is None:
            cls = T.empty((n,), dtype=T.int32, device=conc.device)
        _check(self.L.gckpp_mlp_predict_dev(self.h, n, _ptr(met), _ptr(photol), _ptr(conc),
                                            _ptr(hprev), float(safety), float(hmin), float(hmax),
                                            _ptr(h0), _ptr(cls), _stream()), "gckpp_mlp_predict_dev")
        return h0, cls

    def forward(self, X):
        out = self.torch.empty((X.shape[0], 16), dtype=self.torch.float32, device=X.device)
        _check(self.L.gckpp_mlp_forward_dev(self.h, X.shape[0], _ptr(X), _ptr(out), _stream()),
               "gckpp_mlp_forward_dev")
        return out

    def grad(self, X, yh, ycls, wh=1.0, wc=1.0):
        g = self.torch.empty((MLP_NPARAM,), dtype=self.torch.float32, device=X.device)
        _check(self.L.gckpp_mlp_grad_dev(self.h, X.shape[0], _ptr(X), _ptr(yh), _ptr(ycls),
                                         float(wh), float(wc), _ptr(g), _ptr(self._loss), _stream()),
               "gckpp_mlp_grad_dev")
        return g, self._loss

    def train_step(self, X, yh, ycls, lr, wh=1.0, wc=1.0):
        _check(self.L.gckpp_mlp_train_step_dev(self.h, X.shape[0], _ptr(X), _ptr(yh), _ptr(ycls),
                                               float(wh), float(wc), float(lr), _ptr(self._loss),
                                               _stream()), "gckpp_mlp_train_step_dev")
        return self._loss

    def close(self):
        if self.h:
            self.L.gckpp_mlp_finalize(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
