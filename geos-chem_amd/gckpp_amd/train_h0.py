"""Trains the step-size MLP on step histories logged by the GPU Rosenbrock
solve of synthetic grids and writes geos-chem_amd/models/h0_mlp_<mech>.npz.

Each synthetic grid (72x46x72, gckpp_amd.synth.grid_cells, its own seed and
UTC hour; seeds start at 100 so the benchmark grid, seed 1, is never seen)
gives two logged chemistry steps (dt = 1200 s): its cold step, and the warm
step that follows a synthetic emission/transport perturbation
(gckpp_amd.synth.perturb) with the sun 20 minutes later and Hstart =
KPPHvalue from the cold step.  --grids 24 gives 48 logged steps (~11 M
samples).  The persistent solver logs (H, err) of each cell's first
attempts; labels and classes follow gckpp_amd.learned.

Run on a GPU:  python geos-chem_amd/gckpp_amd/train_h0.py [--grids 24] [--epochs 4]
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
    ap.add_argument("--grids", type=int, default=24, help="synthetic grids (2 logged steps each)")
    ap.add_argument("--grid", default="72x46x72")
    ap.add_argument("--epochs", type=int, default=4)
    ap.add_argument("--batch", type=int, default=8192)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--out", default=None)
    a = ap.parse_args()
    import torch
    from gckpp_amd import hip, learned, mechanism, synth
    from gckpp_amd.options import icntrl_for

    mech = mechanism.load(a.mech)
    nx, ny, nz = (int(v) for v in a.grid.split("x"))
    ic, rc = icntrl_for(a.mech)
    ctx = hip.Context(mech, nx * ny * nz)
    model = learned.StepModel(mech, max_batch=a.batch)
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).cuda()
    Xs, lhs, nss = [], [], []
    t0 = time.time()
    for k in range(a.grids):
        seed, hour = 100 + k, (k * 24.0 / a.grids + 0.5 * (k % 2)) % 24.0
        w0, w1 = learned.warm_pair(mech, nx, ny, nz, seed, hour)
        d0 = dict(met=dev(w0.met), photol=dev(w0.photol), het=dev(w0.het), conc=dev(w0.conc))
        hlog0, n0, e0, c1, h1 = learned.log_step_histories(ctx, d0, ic, rc, w0.dt)
        c1p = synth.perturb(mech, c1.cpu().numpy(), w1.met, seed=seed)
        d1 = dict(met=dev(w1.met), photol=dev(w1.photol), het=dev(w1.het), conc=dev(c1p))
        hlog1, n1, e1, _, _ = learned.log_step_histories(ctx, d1, ic, rc, w0.dt, hstart=h1.clone())
        for d, hlog, nstep, ierr, hprev in ((d0, hlog0, n0, e0, None), (d1, hlog1, n1, e1, h1)):
            ok = ierr == 1
            X = model.raw_features(d["met"], d["photol"], d["conc"], hprev)
            Xs.append(X[torch.from_numpy(ok).cuda()])
            lhs.append(learned.labels(hlog, nstep, w0.dt)[ok])
            nss.append(nstep[ok])
        print("grid %d/%d: seed %d, %.1f UTC: cold step %.1f attempts/cell, warm step %.1f (%.0f s)"
              % (k + 1, a.grids, seed, hour, n0.mean(), n1.mean(), time.time() - t0), flush=True)
        del d0, d1
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
