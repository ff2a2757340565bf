"""Benchmark: fullchem cells/s per 1200 s chemistry step on one MI355X.

Workload: the 4x5-degree 72x46x72 grid (238,464 cells) of synthetic inputs
(gckpp_amd.synth.grid_cells, seed 1), cold start (Hstart = 1e-5 s), dt =
1200 s, repo RTOL/ATOL.  A step is one Do_FullChem pass over every cell
through the C ABI: Update_RCONST, Integrate (Rodas3), the retry of failed
cells and the clipping of negatives.  Inputs are resident in HBM before the
timed region; each step starts from the same initial concentrations.

Prints one JSON line (metric, value, roofline, cpu_baseline, ...).
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT / "geos-chem_amd"))
sys.path.insert(0, str(ROOT))

METRIC = "fullchem cells/sec per 1200 s chem step at repo RTOL: plain GPU vs learned-h0"
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 (vector and matrix) peak
PEAK_HBM_GBS = 8000.0
# FP64 work per Rodas3 attempt of fullchem, counted from the generated code:
# first attempt of a step 145.7 kFLOP, a rejected re-attempt 122.3 kFLOP,
# Update_RCONST ~5 kFLOP per cell.
FLOP_ACC, FLOP_REJ, FLOP_RATES = 145.7e3, 122.3e3, 5.0e3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--grid", default="72x46x72", help="NXxNYxNZ of the synthetic grid")
    p.add_argument("--cpu-sample", type=int, default=8192,
                   help="cells of the same grid timed with the CPU reference (0: skip)")
    p.add_argument("--copy-calibration", action="store_true",
                   help="only run a plain 1 GiB device copy (reference for HBM byte counters)")
    return p.parse_args()


def copy_calibration():
    """A plain device-to-device copy of exactly 1 GiB, five times: the known
    byte count that scales the FETCH_SIZE / WRITE_SIZE counters."""
    import torch
    n = (1 << 30) // 8
    x = torch.ones(n, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": 1 << 30, "repeats": 5}), flush=True)


def cpu_baseline(mech, w, ic, rc, nsample):
    """The compiled reference (oracle/_ref) on a strided sample of the
    workload, all host cores, Do_FullChem semantics (rates, Integrate, retry)."""
    import numpy as np
    from oracle import ref
    if nsample <= 0 or not ref.available():
        return None
    R = ref.RefKPP(mech)
    idx = np.linspace(0, w.ncell - 1, nsample).astype(np.int64)
    ws = w.subset(idx)
    cores = len(os.sched_getaffinity(0))
    R.fullchem_step(ws.subset(np.arange(min(64, nsample))), ic, rc, nthreads=cores)  # warm-up
    t0 = time.perf_counter()
    R.fullchem_step(ws, ic, rc, nthreads=cores)
    dt = time.perf_counter() - t0
    return {"value": round(nsample / dt, 2), "unit": "cells/s", "cores": cores,
            "kind": "reference",
            "sample": "%d cells spread evenly over the same %s grid (every %dth cell), "
                      "reference KPP fullchem core compiled with flang, OpenMP "
                      "DYNAMIC,24 over cells, one chem step" % (nsample, w.name, max(1, w.ncell // nsample))}


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; the chemistry step has no "
              "cross-GPU communication (cells shard independently), so run --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    import numpy as np
    import torch
    from gckpp_amd import hip, mechanism, synth
    from oracle.ref import fullchem_icntrl

    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    mech = mechanism.load("fullchem")
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    w = synth.grid_cells(mech, nx, ny, nz, seed=1)
    n = w.ncell
    ic, rc = fullchem_icntrl()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    met, photol, het, conc0, hstart = dev(w.met), dev(w.photol), dev(w.het), dev(w.conc), dev(w.hstart)
    conc = conc0.clone()
    hnew = torch.empty(n, dtype=torch.float64, device="cuda")
    ist = torch.empty((20, n), dtype=torch.int32, device="cuda")
    ierr = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx = hip.Context(mech, n)

    def step():
        conc.copy_(conc0)
        return ctx.fullchem_step(met, photol, het, conc, hstart, w.dt, ic, rc, hnew=hnew,
                                 istatus=ist, ierr=ierr)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    t_int = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
        t_int.append(ctx.step_timings())
        print("step done: %.1f ms total, integrate %.1f ms" % (t_int[-1][3], t_int[-1][1]),
              file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    cells_per_s = n / (wall / args.steps)

    st = ist.cpu().numpy()
    e = ierr.cpu().numpy()
    nacc, nrej = st[3].astype(np.int64), st[4].astype(np.int64)
    # attempts that are not counted as rejections (before the first accept)
    nstp = st[2].astype(np.int64)
    flop_int = FLOP_ACC * nacc.sum() + FLOP_REJ * (nstp - nacc).sum()
    int_ms = float(np.mean([t[1] for t in t_int]))
    achieved = flop_int / (int_ms * 1e-3) / 1e12
    traffic = None
    tfile = ROOT / "profiles" / "integrate_traffic.json"
    if tfile.exists():
        traffic = json.loads(tfile.read_text()).get("bytes_per_launch")

    out = {
        "metric": METRIC,
        "value": round(cells_per_s, 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "plain GPU Rosenbrock (Rodas3), fullchem, %dx%dx%d grid, "
                               "%d cells, dt=1200 s, cold start, repo RTOL/ATOL" % (nx, ny, nz, n),
                   "cells": n, "solver": "cell-per-lane Rodas3, table-driven"},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
                     "traffic": traffic,
                     "kernel": "k_integrate (Rodas3 integrate, FP64)",
                     "kernel_ms": round(int_ms, 3)},
        "solver_stats": {"mean_steps": float(nstp.mean()), "mean_rejects": float(nrej.mean()),
                         "max_steps": int(nstp.max()), "failed_cells": int((e < 0).sum()),
                         "rates_ms": round(float(np.mean([t[0] for t in t_int])), 3),
                         "retry_ms": round(float(np.mean([t[2] for t in t_int])), 3)},
        "cpu_baseline": cpu_baseline(mech, w, ic, rc, args.cpu_sample),
    }
    print(json.dumps(out), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
