"""Trains the step-size MLP on step histories logged by the GPU Rosenbrock
solve of synthetic grids and writes geos-chem_amd/models/h0_mlp_<mech>.npz.

Each logged step is a cold-start chemistry step (dt = 1200 s) of one
72x46x72 synthetic grid (gckpp_amd.synth.grid_cells) with its own seed and
UTC hour; seeds start at 100 so the benchmark grid (seed 1) is never seen.
The persistent solver logs (H, err) of each cell's first attempts; labels
and classes follow gckpp_amd.learned.

Run on a GPU:  python -m gckpp_amd.train_h0 [--steps 48] [--epochs 4]
(from geos-chem_amd/, or with it on sys.path).
"""
import argparse
import sys
import time
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--mech", default="fullchem")
    ap.add_argument("--steps", type=int, default=48, help="logged grid steps")
    ap.add_argument("--grid", default="72x46x72")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    from gckpp_amd import hip, learned, mechanism, synth
    from oracle.ref import icntrl_for

    mech = mechanism.load(a.mech)
    nx, ny, nz = (int(v) for v in a.grid.split("x"))
    ic, rc = icntrl_for(a.mech)
    ctx = hip.Context(mech, nx * ny * nz)
    model = learned.StepModel(mech, max_batch=a.batch)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    Xs, lhs, nss = [], [], []
    t0 = time.time()
    for k in range(a.steps):
        seed, hour = 100 + k, (k * 24.0 / a.steps + 0.5 * (k % 2)) % 24.0
        w = synth.grid_cells(mech, nx, ny, nz, seed=seed, hour_utc=hour)
        d = dict(met=dev(w.met), photol=dev(w.photol), het=dev(w.het), conc=dev(w.conc))
        hlog, nstep, ierr = learned.log_step_histories(ctx, d, ic, rc, w.dt)
        ok = ierr == 1
        Xs.append(model.raw_features(d["met"], d["photol"], d["conc"])[torch.from_numpy(ok).cuda()])
        lhs.append(learned.labels(hlog, nstep, w.dt)[ok])
        nss.append(nstep[ok])
        print("logged step %d/%d: seed %d, %.1f UTC, %d cells, mean attempts %.1f (%.0f s)"
              % (k + 1, a.steps, seed, hour, int(ok.sum()), nstep.mean(), time.time() - t0), flush=True)
        del d
    Xraw = torch.cat(Xs)
    lnh, nstep = np.concatenate(lhs), np.concatenate(nss)
    del Xs
    model.fit_normalisation(Xraw, lnh, nstep)
    X = model.normalise(Xraw)
    del Xraw
    yh = torch.from_numpy(((lnh - model.out_mu) / model.out_sigma).astype(np.float32)).cuda()
    yc = torch.from_numpy(learned.classes(nstep, model.thresholds)).cuda()
    print("samples %d; ln h label mean %.3f std %.3f; class thresholds %s"
          % (X.shape[0], model.out_mu, model.out_sigma, model.thresholds), flush=True)
    hist = model.train(X, yh, yc, epochs=a.epochs, batch=a.batch, lr=a.lr,
                       log=lambda s: print(s, flush=True))
    out = Path(a.out) if a.out else learned.default_model_path(a.mech)
    model.save(out)
    print("wrote %s (loss per epoch %s, %.0f s)" % (out, ["%.4f" % v for v in hist], time.time() - t0),
          flush=True)


if __name__ == "__main__":
    main()
