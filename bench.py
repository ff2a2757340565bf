"""Benchmark: fullchem cells/s per 1200 s chemistry step on one MI355X,
plain GPU Rosenbrock against the learned-h0 + persistent solver.

Default workload: the 4x5-degree 72x46x72 grid (238,464 cells) of synthetic
inputs (gckpp_amd.synth.grid_cells, seed 1, 12 UTC), cold start, dt = 1200 s,
repo RTOL/ATOL.  A step is one Do_FullChem pass over every cell through the
C ABI: Update_RCONST, Integrate (Rodas3), the retry of failed cells and the
clipping of negatives.  Inputs are resident in HBM before the timed region;
each step starts from the same initial concentrations.

Two legs on the same inputs:
  plain    one lane per cell (k_integrate), Hstart = 1e-5 s (cold start)
  learned  the MLP predicts each cell's h0 and cost class (bf16 MFMA), the
           cells are ordered costliest first, and the persistent solver
           (k_persist, autotuned) integrates them; prediction and ordering
           are inside the timed region.
value = learned-leg cells/s; the plain leg is reported beside it.

Other workloads (--workload): mixed (144x91x72, fullchem/Hg/carbon
70/20/10, one persistent launch per step, 24 steps replayed from a HIP
graph), c90 (6x90x90x72 = 3,499,200 cells in six face-sized chunks).

Prints one JSON line.
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
PEAK_FP64_TFLOPS = 78.6      # MI355X FP64 peak (vector and matrix)
PEAK_HBM_TBPS = 8.0          # MI355X HBM3E peak
# FP64 work per Rodas3 attempt of fullchem, counted from the mechanism tables
# (first attempt of a step 145.7 kFLOP, a rejected re-attempt 122.3 kFLOP).
FLOP_ACC, FLOP_REJ = 145.7e3, 122.3e3


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=3)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--workload", default="grid", choices=["grid", "mixed", "c90"])
    p.add_argument("--grid", default="72x46x72", help="NXxNYxNZ of the synthetic grid")
    p.add_argument("--cpu-sample", type=int, default=8192,
                   help="cells of the same grid timed with the CPU reference (0: skip)")
    p.add_argument("--tight-sample", type=int, default=48,
                   help="cells checked against the RTOL 1e-8 oracle (0: skip)")
    p.add_argument("--safety", type=float, default=1.0, help="factor on the predicted h0")
    p.add_argument("--variant", type=int, default=-1,
                   help="persistent kernel variant (default: autotuned on the workload)")
    p.add_argument("--wpb", type=int, default=1, help="waves per workgroup with --variant")
    p.add_argument("--no-warm", action="store_true", help="skip the warm-step legs")
    p.add_argument("--split", type=int, default=4096,
                   help="warm step: cells (those of the smallest predicted h0) run on the "
                        "cell-per-wave kernel beside the lane kernel, per launch (0: one kernel)")
    p.add_argument("--handoff", type=int, default=98304,
                   help="persistent solver: once at most this many cells of a lane-kernel launch "
                        "are unfinished, they move to the cell-per-wave kernel (0: off)")
    p.add_argument("--no-c90", action="store_true",
                   help="skip the C90-sized leg that follows the grid workload")
    p.add_argument("--no-e2e", action="store_true",
                   help="skip the end-to-end leg (host arrays through gckpp_fullchem_step)")
    p.add_argument("--max-waves", type=int, default=0,
                   help="cap on the waves of a persistent launch (0: every resident wave)")
    p.add_argument("--cpu-box", type=int, default=0, metavar="THREADS",
                   help="only time the compiled CPU reference on the 4,096 synthetic box cells "
                        "(seed 0) with this many OpenMP threads, median of 5 (no GPU needed)")
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(mech, w, ic, rc, nsample, reps=5):
    """The compiled reference (oracle/_ref) on an evenly strided sample of the
    workload, all host cores, Do_FullChem semantics (rates, Integrate, retry,
    clip).  Returns (baseline dict, sample indices, reference outputs)."""
    import numpy as np
    from oracle import ref
    if nsample <= 0 or not ref.available():
        return None, None, None
    R = ref.RefKPP(mech)
    idx = np.linspace(0, w.ncell - 1, nsample).astype(np.int64)
    ws = w.subset(idx)
    cores = len(os.sched_getaffinity(0))
    R.fullchem_step(ws.subset(np.arange(min(64, nsample))), ic, rc, nthreads=cores)  # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        c_ref, ist_ref, _, e_ref = R.fullchem_step(ws, ic, rc, nthreads=cores)
        times.append(time.perf_counter() - t0)
    dt = float(np.median(times))
    base = {"value": round(nsample / dt, 2), "unit": "cells/s", "cores": cores, "kind": "reference",
            "cpu_model": cpu_model(),
            "repetitions": reps, "seconds_per_rep": [round(t, 3) for t in times],
            "sample": "%d cells spread evenly over the same %s grid (every %dth cell), reference "
                      "KPP fullchem core compiled with flang, OpenMP DYNAMIC,24 over cells, one "
                      "chem step; median of %d repetitions (%.2f s each) after a warm-up"
                      % (nsample, w.name, max(1, w.ncell // nsample), reps, dt)}
    return base, idx, (c_ref, ist_ref, e_ref)


def cpu_box(threads, reps=5):
    """The compiled reference's Do_FullChem step on the 4,096 box cells
    (seed 0, cold start) with a stated thread count: the reference's own
    CPU case, timed as the reference runs it (OpenMP DYNAMIC,24)."""
    import numpy as np
    from gckpp_amd import mechanism, options, synth
    from oracle import ref
    m = mechanism.load("fullchem")
    w = synth.box_cells(m, 4096, 0)
    ic, rc = options.fullchem_icntrl()
    R = ref.RefKPP(m)
    R.fullchem_step(w.subset(np.arange(64)), ic, rc, nthreads=threads)   # warm-up
    times = []
    for _ in range(reps):
        t0 = time.perf_counter()
        _, ist, _, e = R.fullchem_step(w, ic, rc, nthreads=threads)
        times.append(time.perf_counter() - t0)
        assert (e == 1).all()
    dt = float(np.median(times))
    print(json.dumps({"metric": "fullchem cells/sec per 1200 s chem step, CPU reference",
                      "value": round(w.ncell / dt, 2), "unit": "cells/s", "threads": threads,
                      "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
                      "seconds_per_rep": [round(t, 3) for t in times],
                      "mean_attempts_per_cell": round(float(ist[2].mean()), 2),
                      "config": {"workload": "4,096 synthetic box cells (seed 0), cold start, "
                                             "dt 1200 s, repo RTOL/ATOL"}}), flush=True)


def cpu_model():
    """The host CPU's model name (first 'model name' of /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


SIGNIFICANT = 1e5   # molec/cm3 (~4e-15 mol/mol at the surface)


def rel_errors(m, c, ref):
    """Scaled error (|c-ref| / (ATOL + RTOL|ref|)) max, and per-species
    relative error over entries above ATOL: max and median; the same scaled
    error over the entries whose reference value exceeds SIGNIFICANT (the
    largest scaled errors of every leg sit in species that are numerically
    zero, e.g. O at night, which the error norm does not control)."""
    import numpy as np
    nv = m.nvar
    a, r = m.atol()[:, None], m.rtol()[:, None]
    sc = np.abs(c[:nv] - ref[:nv]) / (a + r * np.abs(ref[:nv]))
    big = np.abs(ref[:nv]) > a
    rel = np.abs(c[:nv] - ref[:nv])[big] / np.abs(ref[:nv])[big]
    sig = np.abs(ref[:nv]) > SIGNIFICANT
    return {"max_scaled_err": float(sc.max()), "max_rel_err": float(rel.max()) if rel.size else 0.0,
            "median_rel_err": float(np.median(rel)) if rel.size else 0.0,
            "frac_within_rtol": float(np.mean(sc <= 1.0)),
            "max_scaled_err_significant": float(sc[sig].max()) if sig.any() else 0.0,
            "frac_within_rtol_significant": float(np.mean(sc[sig] <= 1.0)) if sig.any() else 1.0}


def parity_entry(err, ncell):
    """Pass/fail of a GPU leg against the compiled reference on the same
    inputs: every variable species of every sampled cell within the repo
    tolerance (scaled error <= 1)."""
    return {"pass": bool(err["max_scaled_err"] <= 1.0), "max_scaled_err": err["max_scaled_err"],
            "cells": int(ncell)}


def reference_step(mech, ws, ic, rc, hstart=None):
    """The compiled reference's chemistry step on the workload ws (per-cell
    Hstart hstart, default ws's own); returns the concentrations."""
    import numpy as np
    from oracle import ref
    if hstart is not None:
        ws.hstart = np.ascontiguousarray(hstart, dtype=np.float64)
    c, _, _, e = ref.RefKPP(mech).fullchem_step(ws, ic, rc, nthreads=len(os.sched_getaffinity(0)))
    return c


def tight_check(mech, w, idx_cells, legs, ic, rc, conc=None):
    """Integrate(0,dt) of a few cells with the C oracle at RTOL 1e-8, ATOL
    1e-6 molec/cm3 (prod/loss dummies excluded), and the errors of each leg
    (legs: name -> result [NSPEC][len(idx_cells)]) against that tight
    solution (the step's clip applied to it), in the repo tolerance's units.
    The repo tolerance bounds each step's error norm (an RMS over species),
    not each species' global error, so a few species stay far from the tight
    solution in every leg; what matters is that the legs are equally close."""
    import numpy as np
    from oracle import orc, ref
    ws = w.subset(idx_cells)
    if conc is not None:          # the cells' state when it is not the workload's own
        ws.conc = np.ascontiguousarray(conc[:, idx_cells])
    rct = ref.RefKPP(mech).rates(ws)
    O = orc.Oracle(mech)
    atol = np.full(mech.nvar, 1e-6)
    atol[mech.atol() >= 1e20] = 1e25
    rtol = np.full(mech.nvar, 1e-8)
    c_t, _, _, e_t = O.integrate(ws.conc, rct, np.zeros(ws.ncell), 0.0, ws.dt, ic, rc, atol=atol,
                                 rtol=rtol, nthreads=len(os.sched_getaffinity(0)))
    ok = e_t == 1
    c_t = np.maximum(c_t, 0.0)
    out = {"cells": int(ok.sum()), "rtol": 1e-8}
    for name, c in legs.items():
        if c is not None:
            out[name] = rel_errors(mech, c[:, ok], c_t[:, ok])
    return out


def grid_bench(args):
    import numpy as np
    import torch
    from gckpp_amd import hip, learned, mechanism, synth
    from gckpp_amd.options import fullchem_icntrl

    mech = mechanism.load("fullchem")
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    t0 = time.time()
    w = synth.grid_cells(mech, nx, ny, nz, seed=1)
    n = w.ncell
    log("workload %s: %d cells (%.1f s)" % (w.name, n, time.time() - t0))
    ic, rc = fullchem_icntrl()
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    met, photol, het, conc0 = dev(w.met), dev(w.photol), dev(w.het), dev(w.conc)
    cold = torch.zeros(n, dtype=torch.float64, device="cuda")
    conc = conc0.clone()
    hnew = torch.empty(n, dtype=torch.float64, device="cuda")
    ist = torch.empty((20, n), dtype=torch.int32, device="cuda")
    ierr = torch.empty(n, dtype=torch.int32, device="cuda")
    ctx = hip.Context(mech, n)

    model = None
    mpath = learned.default_model_path("fullchem")
    if mpath.exists():
        model = learned.StepModel(mech).load(mpath)
    else:
        log("no trained model at %s: the learned leg runs with cold Hstart" % mpath)

    # autotune the persistent solver on the workload itself (a sample smaller
    # than the resident lanes would time a different, latency-only regime),
    # with the tail hand-off it will run with
    ctx.set_persistent_handoff(args.handoff)
    ctx.set_persistent_waves(args.max_waves)
    tune_ms = None
    if args.variant >= 0:
        ctx.set_persistent_config(args.variant, args.wpb)
    else:
        rct = ctx.rates(met, photol, het, conc0)
        tune_ms = ctx.autotune(conc0, rct, cold, 0.0, w.dt, ic, rc)
        del rct
    pcfg = ctx.persistent_config()
    log("autotuned persistent config %s" % pcfg)

    h0 = torch.empty(n, dtype=torch.float64, device="cuda")
    cls = torch.empty(n, dtype=torch.int32, device="cuda")

    def step_plain():
        conc.copy_(conc0)
        ctx.set_solver(hip.SOLVER_PLAIN)
        return ctx.fullchem_step(met, photol, het, conc, cold, w.dt, ic, rc, hnew=hnew, istatus=ist,
                                 ierr=ierr)

    def step_learned():
        conc.copy_(conc0)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        if model is not None:
            model.predict(met, photol, conc, safety=args.safety, dt=w.dt, h0=h0, cls=cls)
            ctx.set_cell_order(hip.order_by_class(cls))
            hs = h0
        else:
            ctx.set_cell_order(None)
            hs = cold
        r = ctx.fullchem_step(met, photol, het, conc, hs, w.dt, ic, rc, hnew=hnew, istatus=ist,
                              ierr=ierr)
        return r

    legs = {}
    for name, fn in (("plain", step_plain), ("learned", step_learned)):
        for _ in range(args.warmup):
            fn()
        torch.cuda.synchronize()
        tim, hos = [], []
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
            tim.append(ctx.step_timings())
            if name == "learned" and args.handoff > 0:
                hos.append(ctx.handoff_stats())
                log("hand-off: %d cells, %d attempts; lane launch %.1f ms, cell-per-wave launch "
                    "%.1f ms" % hos[-1])
            log("%s step done: %.1f ms total, integrate %.1f ms" % (name, tim[-1][3], tim[-1][1]))
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        st = ist.cpu().numpy()
        legs[name] = dict(ms=wall / args.steps * 1e3, cps=n / (wall / args.steps), tim=tim,
                          conc=conc.cpu().numpy(), st=st, ierr=ierr.cpu().numpy(),
                          hnew=hnew.clone(), hos=hos, h0=h0.cpu().numpy())
    ctx.set_cell_order(None)
    L, P = legs["learned"], legs["plain"]
    e2e = end_to_end(args, w, ctx, ic, rc)
    warm = None if args.no_warm else warm_legs(args, mech, w, ctx, model, P, ic, rc)

    # dominant kernel: the persistent integrate (k_persist) of the learned
    # leg; with the hand-off, its own launch (HIP events around it) and its
    # own attempts (all attempts less those of the cell-per-wave launch,
    # counted at the first-attempt cost: rejections are 0.3 % of attempts)
    nacc, nstp = L["st"][3].astype(np.int64), L["st"][2].astype(np.int64)
    flop = FLOP_ACC * nacc.sum() + FLOP_REJ * (nstp - nacc).sum()
    k_ms = float(np.mean([t[1] for t in L["tim"]]))
    handoff = None
    if L["hos"]:
        ho = np.array(L["hos"], dtype=np.float64).mean(axis=0)
        flop -= FLOP_ACC * ho[1]
        k_ms = float(ho[2])
        handoff = {"max_cells": args.handoff, "cells_handed_off": int(round(ho[0])),
                   "cell_per_wave_attempts": int(round(ho[1])), "lane_launch_ms": round(ho[2], 3),
                   "cell_per_wave_launch_ms": round(ho[3], 3),
                   "integrate_ms": round(float(np.mean([t[1] for t in L["tim"]])), 3)}
    achieved = flop / (k_ms * 1e-3) / 1e12
    # HBM bytes per launch from the committed rocprof pass (FETCH_SIZE and
    # WRITE_SIZE, copy-calibrated), when it measured this variant on this grid
    traffic = None
    traffic_src = ("no committed rocprof pass for variant %d, %d cells, hand-off %d"
                   % (pcfg["variant"], n, args.handoff))
    tfile = ROOT / "profiles" / "persist_traffic.json"
    if tfile.exists():
        for tr in json.loads(tfile.read_text()).get("entries", []):
            if (tr.get("workload", "grid") == "grid" and tr.get("variant") == pcfg["variant"] and
                    tr.get("cells") == n and tr.get("handoff", 0) == args.handoff):
                traffic = tr.get("bytes_per_launch")
                traffic_src = tr.get("source", "profiles/k_persist_fullchem_grid72x46x72.md")

    base, sidx, refout = cpu_baseline(mech, w, ic, rc, args.cpu_sample)
    accuracy = {"learned_vs_plain_gpu": rel_errors(mech, L["conc"], P["conc"])}
    parity = {}
    if refout is not None:
        accuracy["learned_vs_reference"] = rel_errors(mech, L["conc"][:, sidx], refout[0])
        accuracy["plain_vs_reference"] = rel_errors(mech, P["conc"][:, sidx], refout[0])
        parity["C2_plain_vs_reference"] = parity_entry(accuracy["plain_vs_reference"], len(sidx))
        if model is not None:
            # the reference started from the learned leg's own h0 (its RCNTRL(3))
            ref_l = reference_step(mech, w.subset(sidx), ic, rc, hstart=L["h0"][sidx])
            accuracy["learned_vs_reference_same_h0"] = rel_errors(mech, L["conc"][:, sidx], ref_l)
            parity["C4_learned_vs_reference_same_h0"] = parity_entry(
                accuracy["learned_vs_reference_same_h0"], len(sidx))
    if args.tight_sample > 0 and sidx is not None:
        pick = np.linspace(0, len(sidx) - 1, args.tight_sample).astype(np.int64)
        cells = sidx[pick]
        accuracy["vs_rtol_1e-8_oracle"] = tight_check(
            mech, w, cells, {"learned": L["conc"][:, cells], "plain_gpu": P["conc"][:, cells],
                             "reference": refout[0][:, pick]}, ic, rc)
        # the cells where the two legs differ most
        nv = mech.nvar
        d = (np.abs(L["conc"][:nv] - P["conc"][:nv]) /
             (mech.atol()[:, None] + mech.rtol()[:, None] * np.abs(P["conc"][:nv]))).max(axis=0)
        worst = np.argsort(d)[::-1][:args.tight_sample // 2]
        accuracy["vs_rtol_1e-8_oracle_most_different_cells"] = tight_check(
            mech, w, worst, {"learned": L["conc"][:, worst], "plain_gpu": P["conc"][:, worst]}, ic, rc)

    out = {
        "metric": METRIC,
        "value": round(L["cps"], 2),
        "unit": "cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(L["ms"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "fullchem chemistry step, %dx%dx%d synthetic grid (seed 1, 12 UTC), "
                               "%d cells, dt=1200 s, cold start, repo RTOL/ATOL; learned h0 (bf16 "
                               "MFMA MLP) + cost-ordered persistent Rodas3" % (nx, ny, nz, n),
                   "cells": n, "learned_model": model is not None, "h0_safety": args.safety,
                   "persistent_config": pcfg, "handoff_max_cells": args.handoff},
        "handoff": handoff,
        "plain_gpu": {"value": round(P["cps"], 2), "unit": "cells/s", "ms_per_step": round(P["ms"], 3),
                      "solver": "one lane per cell (k_integrate), Hstart 1e-5 s"},
        "speedup_vs_plain": round(L["cps"] / P["cps"], 3),
        "roofline": {"bound": "fp64_valu", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 6),
                     "traffic": traffic, "traffic_source": traffic_src,
                     "kernel": "k_persist (persistent Rodas3, FP64 sparse; the algorithmic work "
                               "per attempt from the mechanism tables)",
                     "kernel_ms": round(k_ms, 3),
                     "hbm_TBps": None if traffic is None else round(traffic / (k_ms * 1e-3) / 1e12, 3),
                     "hbm_frac": None if traffic is None else round(traffic / (k_ms * 1e-3) / 1e12 / PEAK_HBM_TBPS, 4)},
        "solver_stats": {
            "learned": {"mean_attempts": float(nstp.mean()), "mean_rejects": float(L["st"][4].mean()),
                        "failed_cells": int((L["ierr"] < 0).sum()),
                        "rates_ms": round(float(np.mean([t[0] for t in L["tim"]])), 3)},
            "plain": {"mean_attempts": float(P["st"][2].mean()), "mean_rejects": float(P["st"][4].mean()),
                      "failed_cells": int((P["ierr"] < 0).sum()),
                      "integrate_ms": round(float(np.mean([t[1] for t in P["tim"]])), 3)}},
        "warm_step": warm,
        "end_to_end": e2e,
        "accuracy": accuracy,
        "parity": parity,
        "autotune_ms": None if tune_ms is None else [[round(float(v), 2) for v in row] for row in tune_ms],
        "cpu_baseline": base,
    }
    ctx.close()
    return out


def end_to_end(args, w, ctx, ic, rc):
    """The same cold step from host arrays through gckpp_fullchem_step (the
    Fortran host's call): the inputs copied to the device, the step, the
    results copied back, all inside the timed region.  Persistent solver
    with the context's hand-off, Hstart = 0 (a host-pointer call takes no
    cell order or predicted h0), pageable host memory; the first call
    allocates the staging buffers and is not timed."""
    import numpy as np
    from gckpp_amd import hip
    if args.no_e2e:
        return None
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ins = [np.ascontiguousarray(a, dtype=np.float64) for a in (w.met, w.photol, w.het)]
    tim = []
    nrep = max(1, min(args.steps, 5))
    for k in range(1 + nrep):      # the first call allocates the staging buffers: untimed
        c = np.ascontiguousarray(w.conc, dtype=np.float64).copy()
        t0 = time.perf_counter()
        hnew, ist, ierr, ok = ctx.fullchem_step_host(*ins, c, None, w.dt, ic, rc)
        tim.append(time.perf_counter() - t0)
    tim = tim[1:]
    wall = float(np.mean(tim))
    nbytes = sum(a.nbytes for a in ins) + 2 * c.nbytes + hnew.nbytes + ist.nbytes + ierr.nbytes
    log("end to end (host arrays): %.1f ms mean of %d calls" % (wall * 1e3, nrep))
    return {"value": round(w.ncell / wall, 2), "unit": "cells/s", "ms_per_step": round(wall * 1e3, 3),
            "calls": nrep, "ms_min": round(min(tim) * 1e3, 3), "ms_max": round(max(tim) * 1e3, 3),
            "host_bytes_moved": int(nbytes), "failed_cells": int((ierr < 0).sum()),
            "path": "host arrays -> gckpp_fullchem_step (staged H2D, persistent solve with the "
                    "hand-off, D2H), Hstart 0, pageable memory"}


def warm_legs(args, mech, w, ctx, model, P, ic, rc):
    """The step after the benchmark step: the plain leg's result, perturbed by
    a synthetic emission/transport operator (gckpp_amd.synth.perturb), with
    the sun 20 minutes later.  Reference semantics: Hstart = KPPHvalue (the
    plain leg's RSTATUS(3)), plain solver; learned: the MLP's h0 from the state
    and KPPHvalue, cost order, persistent solver."""
    import numpy as np
    import torch
    from gckpp_amd import hip, synth
    nx, ny, nz = (int(v) for v in args.grid.split("x"))
    w1 = synth.grid_cells(mech, nx, ny, nz, seed=1, hour_utc=12.0 + w.dt / 3600.0)
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    met, photol, het = dev(w1.met), dev(w1.photol), dev(w1.het)
    conc0 = dev(synth.perturb(mech, P["conc"], w1.met, seed=1))
    kpph = P["hnew"]
    n = w1.ncell
    conc = conc0.clone()
    ist = torch.empty((20, n), dtype=torch.int32, device="cuda")
    ierr = torch.empty(n, dtype=torch.int32, device="cuda")
    hnew = torch.empty(n, dtype=torch.float64, device="cuda")
    h0 = torch.empty(n, dtype=torch.float64, device="cuda")
    cls = torch.empty(n, dtype=torch.int32, device="cuda")

    def plain():
        conc.copy_(conc0)
        ctx.set_solver(hip.SOLVER_PLAIN)
        ctx.set_cell_order(None)
        ctx.fullchem_step(met, photol, het, conc, kpph, w.dt, ic, rc, hnew=hnew, istatus=ist, ierr=ierr)

    def persistent():
        # the persistent solver alone: the reference's Hstart = KPPHvalue, the
        # cells ordered by it (smallest first) with the same split and hand-off
        conc.copy_(conc0)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        if args.split > 0:
            ctx.set_cell_order(hip.order_by_h0(kpph))
            ctx.set_persistent_split(args.split)
        else:
            ctx.set_cell_order(None)
        ctx.fullchem_step(met, photol, het, conc, kpph, w.dt, ic, rc, hnew=hnew, istatus=ist, ierr=ierr)
        ctx.set_persistent_split(0)

    def learned():
        conc.copy_(conc0)
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        hs = kpph
        if model is not None:
            model.predict(met, photol, conc, hprev=kpph, dt=w.dt, h0=h0, cls=cls)
            if args.split > 0:
                # stiffest cells first; the first of them on the cell-per-wave kernel
                ctx.set_cell_order(hip.order_by_h0(h0))
                ctx.set_persistent_split(args.split)
            else:
                ctx.set_cell_order(hip.order_by_class(cls))
            hs = h0
        ctx.fullchem_step(met, photol, het, conc, hs, w.dt, ic, rc, hnew=hnew, istatus=ist, ierr=ierr)
        ctx.set_persistent_split(0)

    out = {}
    res = {}
    for name, fn in (("plain_kpphvalue", plain), ("persistent_kpphvalue", persistent),
                     ("learned", learned)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            fn()
        torch.cuda.synchronize()
        ms = (time.perf_counter() - t0) / args.steps * 1e3
        st = ist.cpu().numpy()
        res[name] = conc.cpu().numpy()
        out[name] = {"value": round(n / ms * 1e3, 2), "unit": "cells/s", "ms_per_step": round(ms, 3),
                     "mean_attempts": float(st[2].mean()), "failed_cells": int((ierr < 0).sum().item())}
        log("warm step %s: %.1f ms, %.2f attempts/cell" % (name, ms, st[2].mean()))
    ctx.set_cell_order(None)
    out["speedup_learned_vs_plain"] = round(out["plain_kpphvalue"]["ms_per_step"] /
                                            out["learned"]["ms_per_step"], 3)
    # what each part contributes: persistence, split and hand-off with the
    # reference's Hstart; then the network's h0 and order on top of them
    out["speedup_persistent_vs_plain"] = round(out["plain_kpphvalue"]["ms_per_step"] /
                                               out["persistent_kpphvalue"]["ms_per_step"], 3)
    out["speedup_learned_vs_persistent"] = round(out["persistent_kpphvalue"]["ms_per_step"] /
                                                 out["learned"]["ms_per_step"], 3)
    out["attempts_saved_frac"] = round(1.0 - out["learned"]["mean_attempts"] /
                                       out["plain_kpphvalue"]["mean_attempts"], 4)
    out["learned_vs_plain"] = rel_errors(mech, res["learned"], res["plain_kpphvalue"])
    if args.tight_sample > 0:
        nv = mech.nvar
        d = (np.abs(res["learned"][:nv] - res["plain_kpphvalue"][:nv]) /
             (mech.atol()[:, None] + mech.rtol()[:, None] * np.abs(res["plain_kpphvalue"][:nv]))).max(axis=0)
        worst = np.argsort(d)[::-1][:args.tight_sample // 2]
        spread = np.linspace(0, n - 1, args.tight_sample // 2).astype(np.int64)
        c0 = conc0.cpu().numpy()
        for key, cells in (("vs_rtol_1e-8_oracle", spread), ("vs_rtol_1e-8_oracle_most_different_cells", worst)):
            out[key] = tight_check(mech, w1, cells, {"learned": res["learned"][:, cells],
                                                     "plain_kpphvalue": res["plain_kpphvalue"][:, cells]},
                                   ic, rc, conc=c0)
    out["workload"] = ("the step after the timed step: its plain-leg result perturbed by synthetic "
                       "emissions (NO, CO, ISOP, CH2O, SO2, ALD2 in P > 700 hPa) and 5% transport "
                       "noise, sun 20 min later, Hstart = KPPHvalue (reference) or learned h0")
    out["learned_split"] = ("cells ordered by learned h0, smallest first; the first %d on the "
                            "cell-per-wave kernel beside the lane kernel" % args.split
                            if args.split > 0 else "one persistent kernel, class order")
    out["handoff_max_cells"] = args.handoff
    return out


def mixed_bench(args):
    """144x91x72 grid, fullchem/Hg/carbon 70/20/10 at random; one persistent
    launch per step over all cells; 24 steps replayed from one HIP graph,
    each step warm-started from the previous step's RSTATUS(3), with a
    synthetic emission/transport perturbation of the state between steps
    (on the device, not timed: without it the closed chemistry settles and
    later steps take one attempt per cell)."""
    import numpy as np
    import torch
    from gckpp_amd import hip, mechanism, synth
    from gckpp_amd.options import icntrl_for
    mechs = [mechanism.load(k) for k in ("fullchem", "Hg", "carbon")]
    t0 = time.time()
    parts, assign = synth.mixed_grid(mechs, 144, 91, 72, seed=3)
    ntot = sum(p.ncell for p in parts)
    log("mixed grid: %s cells (%.1f s)" % ([p.ncell for p in parts], time.time() - t0))
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    ctxs = [hip.Context(m, p.ncell) for m, p in zip(mechs, parts)]
    B = hip.Batch(*ctxs)
    B.set_handoff(args.handoff)
    state = []
    for m, p in zip(mechs, parts):
        ic, rc = icntrl_for(m.name)
        state.append(dict(met=dev(p.met), photol=dev(p.photol), het=dev(p.het), conc0=dev(p.conc),
                          conc=dev(p.conc), hstart=torch.zeros(p.ncell, dtype=torch.float64, device="cuda"),
                          istatus=torch.zeros((20, p.ncell), dtype=torch.int32, device="cuda"),
                          ierr=torch.zeros(p.ncell, dtype=torch.int32, device="cuda"), ic=ic, rc=rc))

    def parts_of():
        return [B.part(met=s["met"], photol=s["photol"], het=s["het"], conc=s["conc"], hstart=s["hstart"],
                       icntrl=s["ic"], rcntrl=s["rc"], istatus=s["istatus"], ierr=s["ierr"]) for s in state]

    def reset():
        for s in state:
            s["conc"].copy_(s["conc0"])
            s["hstart"].zero_()

    # autotune: one cold step per configuration (or the one given)
    tune = {}
    configs = ([(args.variant, args.wpb)] if args.variant >= 0 else
               [(v, g) for v in range(hip.PERSIST_VARIANTS) for g in (1, 2, 4)])
    for v, g in configs:
        try:
            B.set_config(v, g)
        except hip.GckppError:
            continue
        reset()
        torch.cuda.synchronize()
        t = time.perf_counter()
        B.step(parts_of(), 1200.0)
        torch.cuda.synchronize()
        tune[(v, g)] = time.perf_counter() - t
        log("mixed config v%d wpb%d: %.1f ms" % (v, g, tune[(v, g)] * 1e3))
    best = min(tune, key=tune.get)
    B.set_config(*best)
    # wave cap (as on the C90 faces: far more cells than resident lanes):
    # one cold step with and without three waves per SIMD, untimed
    cap_ms = {}
    ncu = torch.cuda.get_device_properties(0).multi_processor_count
    for cap in ((args.max_waves,) if args.max_waves > 0 else (0, 12 * ncu)):
        B.set_waves(cap)
        reset()
        torch.cuda.synchronize()
        t = time.perf_counter()
        B.step(parts_of(), 1200.0)
        torch.cuda.synchronize()
        cap_ms[cap] = (time.perf_counter() - t) * 1e3
        log("mixed wave cap %d: %.1f ms" % (cap, cap_ms[cap]))
    best_cap = min(cap_ms, key=cap_ms.get)
    B.set_waves(best_cap)
    reset()
    B.capture(parts_of(), 1200.0)
    # cells sampled per mechanism (strided) for the first step's accuracy
    sample = [np.linspace(0, p.ncell - 1, min(p.ncell, 2048)).astype(np.int64) for p in parts]
    first = None
    nsteps = 24
    # between two chemistry steps the state moves (emissions, transport):
    # synthetic perturbation on the device, outside the timed replays
    gen = torch.Generator(device="cuda").manual_seed(11)
    wall = 0.0
    attempts = []
    for k in range(nsteps):
        if k > 0:
            for m, s in zip(mechs, state):
                synth.perturb_dev(m, s["conc"], s["met"], gen)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        B.replay()
        torch.cuda.synchronize()
        wall += time.perf_counter() - t0
        attempts.append([float(s["istatus"][2].float().mean().item()) for s in state])
        if k == 0:   # the first (cold) step's result on the sampled cells, untimed
            first = [s["conc"][:, torch.from_numpy(ix).cuda()].cpu().numpy()
                     for s, ix in zip(state, sample)]
        if k % 6 == 5:
            log("replayed %d steps (%.1f s of replays), attempts/cell %s" % (k + 1, wall, attempts[-1]))
    fails = sum(int((s["ierr"] < 0).sum().item()) for s in state)
    # the first replayed step (cold, Hstart 0) against the compiled reference
    # on the same cells, per mechanism
    accuracy, parity = {}, {}
    if args.cpu_sample > 0:
        for m, p, ix, c in zip(mechs, parts, sample, first):
            ic, rc = icntrl_for(m.name)
            cref = reference_step(m, p.subset(ix), ic, rc, hstart=np.zeros(len(ix)))
            err = rel_errors(m, c, cref)
            accuracy["first_step_vs_reference_%s" % m.name] = err
            parity["C3_first_step_vs_reference_%s" % m.name] = parity_entry(err, len(ix))
            log("mixed first step vs reference, %s: max scaled error %.3g on %d cells"
                % (m.name, err["max_scaled_err"], len(ix)))
    out = {
        "metric": "mixed-mechanism cells/sec per 1200 s chem step, 24 graph-replayed steps",
        "value": round(ntot * nsteps / wall, 2), "unit": "cells/s", "n_gpus": 1, "steps": nsteps,
        "warmup": 0, "ms_per_step": round(wall / nsteps * 1e3, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
        "config": {"workload": "144x91x72 grid (seed 3), fullchem/Hg/carbon %s cells assigned at "
                               "random 70/20/10, one persistent launch per step, 24 steps from one "
                               "HIP graph, warm start from RSTATUS(3); between steps (untimed) a "
                               "synthetic emission/transport perturbation of the state" % [p.ncell for p in parts],
                   "persistent_config": {"variant": best[0], "waves_per_group": best[1],
                                         "wave_cap": best_cap},
                   "wave_cap_cold_step_ms": {str(k): round(v, 1) for k, v in cap_ms.items()},
                   "handoff_max_cells": args.handoff,
                   "autotune_first_step_ms": {"v%d_wpb%d" % k: round(v * 1e3, 1) for k, v in tune.items()}},
        "failed_cells_last_step": fails,
        "mean_attempts_per_step": {"first": attempts[0], "last": attempts[-1],
                                   "mean": [float(np.mean([a[q] for a in attempts])) for q in range(3)]},
        "accuracy": accuracy,
        "parity": parity,
    }
    B.close()
    for c in ctxs:
        c.close()
    return out


def _scaled_gpu(mech, c, ref):
    """rel_errors' scaled-error fields of device tensors c vs ref [NSPEC][n]
    (computed on the GPU: the C90 faces are too large to copy for this)."""
    import torch
    nv = mech.nvar
    a = torch.from_numpy(mech.atol()).to(c.device)[:, None]
    r = torch.from_numpy(mech.rtol()).to(c.device)[:, None]
    sc = (c[:nv] - ref[:nv]).abs() / (a + r * ref[:nv].abs())
    sig = ref[:nv].abs() > SIGNIFICANT
    return {"max_scaled_err": float(sc.max().item()), "frac_within_rtol": float((sc <= 1.0).double().mean().item()),
            "max_scaled_err_significant": float(sc[sig].max().item()) if bool(sig.any()) else 0.0,
            "frac_within_rtol_significant": float((sc[sig] <= 1.0).double().mean().item())}


def c90_bench(args):
    """6x90x90x72 = 3,499,200 fullchem cells solved in six face-sized chunks
    (583,200 cells) through one context whose workspace fits HBM.  Two
    steps per leg: the cold step (Hstart = 1e-5 s), then the warm step that
    follows it after a synthetic emission/transport perturbation of the
    state (Hstart = KPPHvalue, or the learned h0 from the state and
    KPPHvalue).  Legs: plain GPU Rosenbrock and learned h0 + persistent.
    value = the learned leg's cells/s over the cold step."""
    import numpy as np
    import torch
    from gckpp_amd import hip, learned, mechanism, synth
    from gckpp_amd.options import fullchem_icntrl
    mech = mechanism.load("fullchem")
    ic, rc = fullchem_icntrl()
    nface = 90 * 90 * 72
    ctx = hip.Context(mech, nface)
    if args.variant >= 0:
        ctx.set_persistent_config(args.variant, args.wpb)
    ctx.set_persistent_handoff(args.handoff)
    ctx.set_persistent_waves(args.max_waves)
    mpath = learned.default_model_path("fullchem")
    model = learned.StepModel(mech).load(mpath) if mpath.exists() else None
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    faces, samples = [], []
    pick = np.linspace(0, nface - 1, 1366).astype(np.int64)   # 8,196 sampled cells over the six faces
    for f in range(6):
        w = synth.grid_cells(mech, 90, 90, 72, seed=200 + f, hour_utc=(12.0 + 4.0 * f) % 24.0)
        faces.append(dict(met=dev(w.met), photol=dev(w.photol), het=dev(w.het), conc0=dev(w.conc)))
        samples.append(w.subset(pick))
        log("face %d generated" % f)
    del w
    h0 = torch.empty(nface, dtype=torch.float64, device="cuda")
    cls = torch.empty(nface, dtype=torch.int32, device="cuda")
    cold = torch.zeros(nface, dtype=torch.float64, device="cuda")
    conc = torch.empty_like(faces[0]["conc0"])
    gen = torch.Generator(device="cuda")

    def run(use_learned, warm):
        """One step of every face; returns wall s, kernel ms, attempts, ..."""
        fails, att, acc_att, kern_ms = 0, [], 0, 0.0
        ho = [0, 0, 0.0, 0.0]   # hand-off: cells, cell-per-wave attempts, lane ms, cell-per-wave ms
        ctx.set_solver(hip.SOLVER_PERSISTENT if use_learned else hip.SOLVER_PLAIN)
        outs = []
        wall = 0.0
        for k, f in enumerate(faces):
            src = f["warm_in"] if warm else f["conc0"]
            conc.copy_(src)
            hprev = f["kpph"] if warm else None
            hs = hprev if warm else cold
            ctx.set_cell_order(None)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if use_learned and model is not None:
                model.predict(f["met"], f["photol"], conc, hprev=hprev, safety=args.safety, h0=h0, cls=cls)
                if warm and args.split > 0:
                    ctx.set_cell_order(hip.order_by_h0(h0))
                    ctx.set_persistent_split(args.split)
                else:
                    ctx.set_cell_order(hip.order_by_class(cls))
                hs = h0
            hnew, ist, ierr, ok = ctx.fullchem_step(f["met"], f["photol"], f["het"], conc, hs, 1200.0,
                                                    ic, rc)
            ctx.set_persistent_split(0)
            torch.cuda.synchronize()
            wall += time.perf_counter() - t0
            kern_ms += float(ctx.step_timings()[1])
            if use_learned and args.handoff > 0:
                ho = [a + b for a, b in zip(ho, ctx.handoff_stats())]
            fails += int((ierr < 0).sum().item())
            att.append(float(ist[2].float().mean().item()))
            acc_att += int(ist[3].sum().item())
            tot_att = int(ist[2].sum().item())
            outs.append((conc.clone(), hnew.clone(), acc_att, tot_att,
                         hs[pick].cpu().numpy() if hs is not None else None))
            acc_att = 0
        ctx.set_cell_order(None)
        log("c90 %s %s step: %.1f s" % ("learned + persistent" if use_learned else "plain",
                                         "warm" if warm else "cold", wall))
        return wall, kern_ms, fails, att, outs, ho

    # wave cap of the persistent launches (untimed, on the first face's cold
    # step): a face has more than twice the lanes the GPU holds at once, and
    # in that steady state of work stealing fewer resident waves (three per
    # SIMD instead of every one that fits) can stream the slots faster
    cap_ms = {}
    if args.max_waves == 0 and model is not None:
        ncu = torch.cuda.get_device_properties(0).multi_processor_count
        f = faces[0]
        ctx.set_solver(hip.SOLVER_PERSISTENT)
        for cap in (0, 12 * ncu):
            ctx.set_persistent_waves(cap)
            conc.copy_(f["conc0"])
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            model.predict(f["met"], f["photol"], conc, safety=args.safety, h0=h0, cls=cls)
            ctx.set_cell_order(hip.order_by_class(cls))
            ctx.fullchem_step(f["met"], f["photol"], f["het"], conc, h0, 1200.0, ic, rc)
            torch.cuda.synchronize()
            cap_ms[cap] = (time.perf_counter() - t0) * 1e3
            ctx.set_cell_order(None)
        best_cap = min(cap_ms, key=cap_ms.get)
        ctx.set_persistent_waves(best_cap)
        log("c90 wave cap: %s ms -> %d" % ({k: round(v, 1) for k, v in cap_ms.items()}, best_cap))
    res = {}
    for warm in (False, True):
        if warm:
            # the warm step's input: the plain leg's cold result, perturbed
            for f, (c, hn, _, _, _) in zip(faces, res[("plain", False)][4]):
                f["warm_in"] = c.clone()
                synth.perturb_dev(mech, f["warm_in"], f["met"], gen.manual_seed(7))
                f["kpph"] = hn
        for leg in ("learned", "plain"):
            res[(leg, warm)] = run(leg == "learned", warm)
    n = 6 * nface

    def leg_json(key):
        wall, kms, fails, att, outs, ho = res[key]
        return {"value": round(n / wall, 2), "unit": "cells/s", "ms_per_step": round(wall * 1e3, 3),
                "integrate_kernel_ms": round(kms, 3), "failed_cells": fails,
                "mean_attempts_per_face": [round(a, 3) for a in att]}

    # accuracy: learned vs plain on every cell (GPU), and both against the
    # CPU reference and the RTOL 1e-8 oracle on samples of the cold step
    acc = {}
    for warm in (False, True):
        errs = [_scaled_gpu(mech, a[0], b[0]) for a, b in zip(res[("learned", warm)][4], res[("plain", warm)][4])]
        acc["learned_vs_plain_gpu_" + ("warm" if warm else "cold")] = {
            "max_scaled_err": max(e["max_scaled_err"] for e in errs),
            "frac_within_rtol": float(np.mean([e["frac_within_rtol"] for e in errs])),
            "max_scaled_err_significant": max(e["max_scaled_err_significant"] for e in errs),
            "frac_within_rtol_significant": float(np.mean([e["frac_within_rtol_significant"] for e in errs]))}
    ws = samples[0]
    for k in range(1, 6):
        o = samples[k]
        ws = synth.Workload(mech, np.concatenate([ws.met, o.met], 1), np.concatenate([ws.photol, o.photol], 1),
                            np.concatenate([ws.het, o.het], 1) if ws.het.shape[1] > 1 else ws.het,
                            np.concatenate([ws.conc, o.conc], 1), np.concatenate([ws.hstart, o.hstart]), ws.dt,
                            "c90 faces")
    gl = np.concatenate([res[("learned", False)][4][k][0][:, pick].cpu().numpy() for k in range(6)], 1)
    gp = np.concatenate([res[("plain", False)][4][k][0][:, pick].cpu().numpy() for k in range(6)], 1)
    base, sidx, refout = cpu_baseline(mech, ws, ic, rc, ws.ncell)
    parity = {}
    if refout is not None:
        acc["learned_vs_reference_cold"] = rel_errors(mech, gl[:, sidx], refout[0])
        acc["plain_vs_reference_cold"] = rel_errors(mech, gp[:, sidx], refout[0])
        parity["C5_plain_vs_reference"] = parity_entry(acc["plain_vs_reference_cold"], len(sidx))
        if model is not None:
            hl = np.concatenate([res[("learned", False)][4][k][4] for k in range(6)])
            ref_l = reference_step(mech, ws.subset(sidx), ic, rc, hstart=hl[sidx])
            acc["learned_vs_reference_same_h0_cold"] = rel_errors(mech, gl[:, sidx], ref_l)
            parity["C5_learned_vs_reference_same_h0"] = parity_entry(
                acc["learned_vs_reference_same_h0_cold"], len(sidx))
        if args.tight_sample > 0:
            tp = np.linspace(0, len(sidx) - 1, args.tight_sample).astype(np.int64)
            cells = sidx[tp]
            acc["vs_rtol_1e-8_oracle_cold"] = tight_check(
                mech, ws, cells, {"learned": gl[:, cells], "plain_gpu": gp[:, cells],
                                  "reference": refout[0][:, tp]}, ic, rc)
    # roofline of the learned leg's integrate kernels over the cold step
    flop = 0.0
    for (_, _, nacc, ntot, _) in res[("learned", False)][4]:
        flop += FLOP_ACC * nacc + FLOP_REJ * (ntot - nacc)
    kms = res[("learned", False)][1]
    ho = res[("learned", False)][5]
    if args.handoff > 0:   # k_persist's own launches and attempts (see grid_bench)
        flop -= FLOP_ACC * ho[1]
        kms = ho[2]
    achieved = flop / (kms * 1e-3) / 1e12
    # HBM bytes per chunk launch from the committed rocprof passes over this
    # workload (FETCH_SIZE and WRITE_SIZE, copy-calibrated, the mean of the
    # six lane launches), when they measured this configuration
    traffic = None
    pc = ctx.persistent_config()
    traffic_src = ("no committed rocprof pass for variant %d, %d waves, hand-off %d"
                   % (pc["variant"], pc["waves"], args.handoff))
    tfile = ROOT / "profiles" / "persist_traffic.json"
    if tfile.exists():
        for tr in json.loads(tfile.read_text()).get("entries", []):
            if (tr.get("workload") == "c90" and tr.get("variant") == pc["variant"] and
                    tr.get("cells") == nface and tr.get("handoff", 0) == args.handoff and
                    tr.get("waves") == pc["waves"]):
                traffic = tr.get("bytes_per_launch")
                traffic_src = tr.get("source")
    hbm = None if traffic is None else 6 * traffic / (kms * 1e-3) / 1e12
    L, P = leg_json(("learned", False)), leg_json(("plain", False))
    Lw, Pw = leg_json(("learned", True)), leg_json(("plain", True))
    return {"metric": METRIC + " (C90-sized)", "value": L["value"], "unit": "cells/s", "n_gpus": 1,
            "steps": 1, "warmup": 0, "ms_per_step": L["ms_per_step"], "higher_is_better": True,
            "scaling": "weak", "vs_baseline": None, "dtype": "f64", "data": "synthetic",
            "config": {"workload": "6x90x90x72 = 3,499,200 fullchem cells (six 90x90x72 synthetic "
                                   "faces, seeds 200-205), six 583,200-cell chunks; cold step, then "
                                   "the warm step after a synthetic emission/transport perturbation; "
                                   "learned h0 + persistent against plain GPU",
                       "learned_model": model is not None,
                       "persistent_config": ctx.persistent_config(),
                       "wave_cap_first_face_cold_ms": {str(k): round(v, 1) for k, v in cap_ms.items()},
                       "handoff_max_cells": args.handoff},
            "handoff_cold": None if args.handoff <= 0 else {
                "cells_handed_off": ho[0], "cell_per_wave_attempts": ho[1],
                "lane_launch_ms": round(ho[2], 3), "cell_per_wave_launch_ms": round(ho[3], 3)},
            "learned": L, "plain_gpu": P, "speedup_vs_plain": round(P["ms_per_step"] / L["ms_per_step"], 3),
            "warm_step": {"learned": Lw, "plain_kpphvalue": Pw,
                          "speedup_vs_plain": round(Pw["ms_per_step"] / Lw["ms_per_step"], 3),
                          "learned_split": ("cells ordered by learned h0, smallest first; the first "
                                            "%d of each chunk on the cell-per-wave kernel beside "
                                            "the lane kernel" % args.split
                                            if args.split > 0 else "one persistent kernel")},
            "roofline": {"bound": "fp64_valu", "achieved": round(achieved, 4), "peak": PEAK_FP64_TFLOPS,
                         "unit": "TFLOP/s", "frac": round(achieved / PEAK_FP64_TFLOPS, 6), "traffic": traffic,
                         "traffic_source": traffic_src,
                         "kernel": "k_persist over the six chunks of the cold step (traffic: per chunk launch)",
                         "kernel_ms": round(kms, 3),
                         "hbm_TBps": None if hbm is None else round(hbm, 3),
                         "hbm_frac": None if hbm is None else round(hbm / PEAK_HBM_TBPS, 4)},
            "accuracy": acc, "parity": parity, "cpu_baseline": base}


def main():
    args = parse()
    if args.copy_calibration:
        return copy_calibration()
    if args.cpu_box:
        return cpu_box(args.cpu_box)
    if args.gpus != 1:
        print("bench.py measures one GPU: the chemistry step has no cross-GPU communication "
              "(cells shard independently), so run it with --gpus 1", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs a HIP device"
    out = {"grid": grid_bench, "mixed": mixed_bench, "c90": c90_bench}[args.workload](args)
    if args.workload == "grid" and not args.no_c90:
        # the C90-sized step (3,499,200 cells in six chunks) beside the grid's line:
        # one cold and one warm step of each leg, with its own roofline,
        # accuracy, parity and CPU baseline
        torch.cuda.empty_cache()
        c90 = c90_bench(args)
        out["c90"] = {k: c90[k] for k in ("value", "unit", "ms_per_step", "config", "handoff_cold",
                                          "learned", "plain_gpu", "speedup_vs_plain", "warm_step",
                                          "roofline", "accuracy", "cpu_baseline")}
        out["parity"].update(c90["parity"])
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
