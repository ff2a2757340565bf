"""GPU Update_RCONST against the compiled reference's RCONST (golden
fixtures): clear-sky box cells, and cells with rich per-cell HetState
records that drive every heterogeneous rate law through its branches.

Tolerance: each rate is a short chain of FP64 operations and transcendental
calls (exp, log10, pow, sqrt) whose device and host implementations may
differ by an ulp or two, so gas-phase and clear-sky rates agree to 1e-12
relative.  The cloud-uptake laws (CloudHet's entrainment-limited rate, the
reacto-diffusive coth(x)-1/x correction) subtract nearly equal terms, which
turns ulp-level differences into up to ~1e-8 relative; for rich HetState
records the bound is 1e-7.  Exact zeros (rates a branch switches off) must be
exact zeros on both sides.
"""
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
GOLD = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def ctx(fullchem):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    from gckpp_amd import hip
    c = hip.Context(fullchem, 4096)
    yield c
    c.close()


def _gpu_rates(ctx, g):
    import torch
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    return ctx.rates(d(g["met"]), d(g["photol"]), d(g["het"]), d(g["conc"])).cpu().numpy()


def _compare(got, ref, fullchem):
    assert np.isfinite(got).all()
    assert np.array_equal(got == 0.0, ref == 0.0), np.argwhere((got == 0.0) != (ref == 0.0))[:5]
    nz = ref != 0.0
    rel = np.zeros_like(ref)
    rel[nz] = np.abs(got[nz] - ref[nz]) / np.abs(ref[nz])
    worst = np.unravel_index(rel.argmax(), rel.shape)
    print("max rel err %.3g at reaction %d (%s)" % (rel.max(), worst[0],
          (fullchem.d["eqn_names"] or [""] * 9999)[worst[0]].strip()))
    return rel


def test_rates_clear_sky_box_cells(ctx, fullchem):
    print("case: rates on golden clear-sky box cells", flush=True)
    g = dict(np.load(GOLD / "fullchem_box.npz"))
    rel = _compare(_gpu_rates(ctx, g), g["rconst"], fullchem)
    assert rel.max() < 1e-12


def test_rates_rich_het_state(ctx, fullchem):
    print("case: rates with rich per-cell HetState", flush=True)
    g = dict(np.load(GOLD / "fullchem_het.npz"))
    rel = _compare(_gpu_rates(ctx, g), g["rconst"], fullchem)
    assert rel.max() < 1e-7
    # laws without cloud terms stay at ulp level: every rate up to the first
    # heterogeneous reaction (RCONST 781, HO2 uptake)
    assert rel[:780].max() < 1e-12
