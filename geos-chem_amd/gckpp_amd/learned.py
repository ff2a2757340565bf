"""Learned initial step size and stiffness class for the Rosenbrock solve.

Each cell's integration starts at Hstart: 1e-5 s on a cold step
(KPP/fullchem/gckpp_Integrator.F90:389-395, RCNTRL(3)=0), else the previous
step's last proposed step, KPPHvalue = RSTATUS(3)
(GeosCore/fullchem_mod.F90:1016-1020,1294).  After the operators that run
between chemistry steps (emissions, transport) a cell's state has moved and
that step can be far off, costing rejected attempts; on a cold step the
controller needs several accepted steps before the error limits the step.
The MLP (csrc/mlp.hip) predicts, from the cell's state, the step size the
controller settles on and a cost class; the solve then starts each cell at
the predicted size and the persistent solver takes the costliest cells
first.  The controller is unchanged: it accepts or rejects every step
exactly as before, so results meet the same tolerance.

Training data are the step histories the GPU Rosenbrock logs
(gckpp_set_step_log): (H, err) of each cell's first attempts, on cold steps
of synthetic grids and on the warm steps that follow them after a synthetic
emission/transport perturbation (gckpp_amd.synth.perturb).  Labels: see
labels().  Class: quartile of the cell's number of attempts (ISTATUS(3)) in
the training set, 0 cheapest .. 3 costliest.
"""
import json
from pathlib import Path

import numpy as np

# PHOTOL rows (0-based slots) and species of the inputs
J_SLOTS = [1, 2, 10, 8, 9, 6, 7, 15, 14, 11, 13, 60]
SPECIES = ["O3", "NO", "NO2", "OH", "HO2", "CO", "CH4", "ISOP", "HNO3", "H2O2", "CH2O", "NO3",
           "N2O5", "PAN", "SO2", "BrO", "Cl", "ClO", "MP", "ALD2", "ACET", "C2H6", "MO2", "HNO2"]
NLOG = 24                       # attempts logged per cell
NIN = 42                        # model inputs (csrc/mlp.h)
FAC_MIN, FAC_MAX, FAC_SAFE = 0.2, 6.0, 0.9
MODEL_DIR = Path(__file__).resolve().parent.parent / "models"


def inputs(mech):
    """(jslot, spc) index arrays of the model inputs for a mechanism."""
    return (np.array(J_SLOTS, dtype=np.int32),
            np.array([mech.species.index(s) for s in SPECIES], dtype=np.int32))


def labels(hlog, nstep, dt):
    """ln(h) label per cell from a step log hlog[2*NLOG][ncell] (float) and
    the number of attempts per cell nstep[ncell] (ISTATUS(3)).  If the first
    attempt was rejected (err > 1): the step of the first accepted attempt.
    Otherwise: at the first error-limited attempt (err > (FacSafe/FacMax)^3),
    the step the controller proposes; if none is logged, the last H times
    FacMax, at most dt."""
    nlog = hlog.shape[0] // 2
    H = hlog[0::2].astype(np.float64)
    E = hlog[1::2].astype(np.float64)
    n = hlog.shape[1]
    cols = np.arange(n)
    valid = np.arange(nlog)[:, None] < np.minimum(nstep, nlog)[None, :]
    limited = valid & (E > (FAC_SAFE / FAC_MAX) ** 3)
    first = np.where(limited.any(axis=0), limited.argmax(axis=0), -1)
    k = np.maximum(first, 0)
    fac = np.minimum(FAC_MAX, np.maximum(FAC_MIN, FAC_SAFE / np.maximum(E[k, cols], 1e-300) ** (1 / 3)))
    h_lim = H[k, cols] * fac
    last = np.maximum(np.minimum(nstep, nlog) - 1, 0)
    h_free = np.minimum(H[last, cols] * FAC_MAX, dt)
    h = np.where(first >= 0, h_lim, h_free)
    accepted = valid & (E <= 1.0)
    first_acc = np.where(accepted.any(axis=0), accepted.argmax(axis=0), 0)
    rejected0 = valid[0] & (E[0] > 1.0)
    h = np.where(rejected0, H[first_acc, cols], h)
    return np.log(np.clip(h, 1e-12, dt))


def classes(nstep, thresholds):
    return np.searchsorted(np.asarray(thresholds), nstep, side="right").astype(np.int32)


class StepModel:
    """The MLP on the GPU plus its normalisation and class thresholds."""

    def __init__(self, mech, max_batch=65536):
        from . import hip
        self.mech = mech
        self.jslot, self.spc = inputs(mech)
        self.net = hip.MLP(max_batch, self.jslot, self.spc)
        self.mean = np.zeros(48, np.float32)
        self.std = np.ones(48, np.float32)
        self.out_mu, self.out_sigma = 0.0, 1.0
        self.thresholds = [0, 0, 0]

    # -- persistence ----------------------------------------------------------
    def save(self, path):
        path = Path(path)
        path.parent.mkdir(parents=True, exist_ok=True)
        np.savez_compressed(path, params=self.net.params(), mean=self.mean, std=self.std,
                            out=np.array([self.out_mu, self.out_sigma]),
                            thresholds=np.asarray(self.thresholds, dtype=np.float64),
                            jslot=self.jslot, spc=self.spc)

    def load(self, path):
        z = np.load(path)            # arrays only (allow_pickle stays False)
        assert np.array_equal(z["jslot"], self.jslot) and np.array_equal(z["spc"], self.spc)
        self.mean, self.std = z["mean"].astype(np.float32), z["std"].astype(np.float32)
        self.out_mu, self.out_sigma = (float(v) for v in z["out"])
        self.thresholds = list(z["thresholds"])
        self.net.set_norm(self.mean, self.std, self.out_mu, self.out_sigma)
        self.net.set_params(z["params"])
        return self

    # -- data -----------------------------------------------------------------
    def raw_features(self, met, photol, conc, hprev=None):
        """[ncell][48] un-normalised inputs, computed by the features kernel."""
        self.net.set_norm(np.zeros(48, np.float32), np.ones(48, np.float32), 0.0, 1.0)
        X = self.net.features(met, photol, conc, hprev)
        self.net.set_norm(self.mean, self.std, self.out_mu, self.out_sigma)
        return X

    def fit_normalisation(self, Xraw, lnh, nstep):
        """Input mean/std, label mean/std and class thresholds from training data."""
        x = Xraw[:, :NIN].double()
        self.mean[:NIN] = x.mean(dim=0).cpu().numpy()
        self.std[:NIN] = np.maximum(x.std(dim=0).cpu().numpy(), 1e-6)
        self.out_mu, self.out_sigma = float(np.mean(lnh)), float(max(np.std(lnh), 1e-6))
        self.thresholds = [float(v) for v in np.quantile(nstep, [0.25, 0.5, 0.75])]
        self.net.set_norm(self.mean, self.std, self.out_mu, self.out_sigma)

    def normalise(self, Xraw):
        import torch
        m = torch.from_numpy(self.mean).to(Xraw.device)
        s = torch.from_numpy(self.std).to(Xraw.device)
        X = (Xraw - m) / s
        X[:, NIN:] = 0.0
        return X.contiguous()

    # -- training -------------------------------------------------------------
    def train(self, X, yh, yc, epochs=4, batch=4096, lr=2e-3, wh=1.0, wc=0.5, seed=0, log=None):
        """Adam on shuffled mini-batches with the GPU training kernels; X, yh
        (normalised ln h), yc device tensors.  Returns the loss per epoch."""
        import torch
        n = X.shape[0]
        g = torch.Generator(device="cpu").manual_seed(seed)
        rng = np.random.default_rng(seed)
        p = rng.normal(0.0, 1.0, 41360).astype(np.float32)
        # He-style initial scale per layer, zero biases
        offs = [(0, 48, 128), (6272, 128, 128), (22784, 128, 128), (39296, 128, 16)]
        q = np.zeros_like(p)
        for o, a, b in offs:
            q[o:o + a * b] = p[o:o + a * b] * np.sqrt(2.0 / a)
        self.net.set_params(q)
        hist = []
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).to(X.device)
            tot, nb = 0.0, 0
            lr_ep = lr * (0.5 ** ep)
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                loss = self.net.train_step(X[idx].contiguous(), yh[idx].contiguous(),
                                           yc[idx].contiguous(), lr_ep, wh, wc)
                if nb % 50 == 0:
                    tot += float(loss.item())
                nb += 1
            hist.append(tot / max(1, (nb + 49) // 50))
            if log:
                log("epoch %d: loss %.4f" % (ep, hist[-1]))
        return hist

    # -- use --------------------------------------------------------------------
    def predict(self, met, photol, conc, hprev=None, safety=1.0, dt=1200.0, h0=None, cls=None):
        """Hstart per cell (hprev: the previous step size per cell, None on a
        cold step) and the cost class."""
        return self.net.predict(met, photol, conc, hprev=hprev, safety=safety, hmin=1e-5, hmax=dt,
                                h0=h0, cls=cls)


def log_step_histories(ctx, w_dev, ic, rc, dt, nlog=NLOG, hstart=None):
    """One persistent chemistry step of a workload with the step log on,
    from Hstart = hstart (device tensor; None: cold start).  Returns (hlog
    [2*nlog][n] float32 host, nstep [n] host, ierr host, conc after the step,
    RSTATUS(3)) -- the last two device tensors."""
    import torch
    from . import hip
    n = w_dev["conc"].shape[1]
    hlog = torch.zeros((2 * nlog, n), dtype=torch.float32, device="cuda")
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_step_log(hlog, nlog)
    conc = w_dev["conc"].clone()
    hs = torch.zeros(n, dtype=torch.float64, device="cuda") if hstart is None else hstart
    hnew, ist, ierr, ok = ctx.fullchem_step(w_dev["met"], w_dev["photol"], w_dev["het"], conc, hs,
                                            dt, ic, rc)
    ctx.set_step_log(None)
    return hlog.cpu().numpy(), ist[2].cpu().numpy(), ierr.cpu().numpy(), conc, hnew


def warm_pair(mech, nx, ny, nz, seed, hour, dt=1200.0):
    """A grid for a warm chemistry step: the workload at `hour` (its cold
    step gives the state and KPPHvalue), and the met/photolysis 20 minutes
    later that the next step sees; the caller perturbs the state between."""
    from . import synth
    w0 = synth.grid_cells(mech, nx, ny, nz, seed=seed, hour_utc=hour, dt=dt)
    w1 = synth.grid_cells(mech, nx, ny, nz, seed=seed, hour_utc=(hour + dt / 3600.0) % 24.0, dt=dt)
    return w0, w1


def default_model_path(mech_name="fullchem"):
    return MODEL_DIR / ("h0_mlp_%s.npz" % mech_name)
