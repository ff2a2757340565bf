"""gckpp_integrate_noalloc_dev, through the C ABI, against gckpp_integrate_dev
with the persistent solver's cell-per-wave variant on the same context and
inputs: bitwise the same concentrations, ISTATUS, RSTATUS and IERR, on two
calls in a row (the second with fewer cells, so it runs on the scratch the
first left behind), for fullchem and Hg; the device's free memory is the
same before and after the call; a wave cap beyond the static scratch is
refused with a message."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def T():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _noalloc(ctx, conc, rct, hs, tin, tout, ic, rc, ist, rst, ierr):
    from gckpp_amd import hip
    ic = np.ascontiguousarray(ic, dtype=np.int32)
    rc = np.ascontiguousarray(rc, dtype=np.float64)
    r = ctx.L.gckpp_integrate_noalloc_dev(ctx.h, conc.shape[1], float(tin), float(tout),
                                          ctypes.c_void_p(ic.ctypes.data),
                                          ctypes.c_void_p(rc.ctypes.data), hip._ptr(hs),
                                          hip._ptr(conc), hip._ptr(rct), hip._ptr(ist),
                                          hip._ptr(rst), hip._ptr(ierr), hip._stream())
    if r != 0:
        raise hip.GckppError(ctx.L.gckpp_last_error().decode())


@pytest.mark.parametrize("name", ["fullchem", "Hg"])
def test_noalloc_equals_integrate_dev(name, mech_of, T):
    from gckpp_amd import hip, options, synth
    m = mech_of(name)
    ic, rc = options.icntrl_for(name)
    ctx = hip.Context(m, 2048)
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_persistent_config(hip.WAVE_VARIANT)
    dev = lambda a: T.from_numpy(np.ascontiguousarray(a)).cuda()
    for n, seed in ((1500, 5), (300, 6)):
        print("case: no-allocation Integrate, %s, %d cells" % (name, n), flush=True)
        w = synth.box_cells(m, 4096, seed).subset(np.r_[0:n // 2, 2048:2048 + n - n // 2])
        w.hstart = np.where(np.arange(n) % 3 == 0, 0.0, 20.0)
        rct = ctx.rates(dev(w.met), dev(w.photol), dev(w.het), dev(w.conc))
        hs = dev(w.hstart)
        c1 = dev(w.conc)
        ist1, rst1, e1 = ctx.integrate(c1, rct, hs, 0.0, w.dt, ic, rc)
        c2 = dev(w.conc)
        ist2 = T.full((20, n), -99, dtype=T.int32, device="cuda")
        rst2 = T.full((20, n), -99.0, dtype=T.float64, device="cuda")
        e2 = T.full((n,), -99, dtype=T.int32, device="cuda")
        T.cuda.synchronize()
        free0 = T.cuda.mem_get_info()[0]
        _noalloc(ctx, c2, rct, hs, 0.0, w.dt, ic, rc, ist2, rst2, e2)
        T.cuda.synchronize()
        free1 = T.cuda.mem_get_info()[0]
        assert free1 >= free0, "the call allocated device memory"
        assert (e1.cpu().numpy() == 1).all()
        assert T.equal(c1, c2) and T.equal(ist1, ist2) and T.equal(rst1, rst2) and T.equal(e1, e2)
    # a wave cap the static scratch cannot hold is refused
    ctx.set_persistent_waves(1 << 20)
    with pytest.raises(hip.GckppError, match="exceeds"):
        _noalloc(ctx, c2, rct, hs, 0.0, w.dt, ic, rc, ist2, rst2, e2)
    ctx.set_persistent_waves(0)
    ctx.close()
