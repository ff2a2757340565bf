"""Learned initial step size and stiffness class for the Rosenbrock solve.

A cold chemistry step starts every cell at Hstart = 1e-5 s
(KPP/fullchem/gckpp_Integrator.F90:389-395, RCNTRL(3)=0) and the step-size
controller needs several accepted steps, each at most FacMax=6 times the
last, before the error estimate limits the step.  The MLP (csrc/mlp.hip)
predicts, from the cell's state, the step size the controller settles on
and a cost class; the solve then starts each cell at the predicted size and
the persistent solver takes the costliest cells first.  The controller is
unchanged: it accepts or rejects every step exactly as before, so results
meet the same tolerance.

Training data are the step histories the GPU Rosenbrock logs
(gckpp_set_step_log): (H, err) of each cell's first attempts.  Label of a
cell: at its first attempt whose error limits the step
(err > (FacSafe/FacMax)^3), the step the controller proposes,
H*min(FacMax, max(FacMin, FacSafe/err^(1/3))) (gckpp_Integrator.F90:725-735);
if no logged attempt is error-limited, the last H times FacMax, at most dt.
Class: quartile of the cell's number of attempts (ISTATUS(3)) in the
training set, 0 cheapest .. 3 costliest.
"""
import json
from pathlib import Path

import numpy as np

# PHOTOL rows (0-based slots) and species of the inputs
J_SLOTS = [1, 2, 10, 8, 9, 6, 7, 15, 14, 11, 13, 60]
SPECIES = ["O3", "NO", "NO2", "OH", "HO2", "CO", "CH4", "ISOP", "HNO3", "H2O2", "CH2O", "NO3",
           "N2O5", "PAN", "SO2", "BrO", "Cl", "ClO", "MP", "ALD2", "ACET", "C2H6", "MO2", "HNO2"]
NLOG = 24                       # attempts logged per cell
FAC_MIN, FAC_MAX, FAC_SAFE = 0.2, 6.0, 0.9
MODEL_DIR = Path(__file__).resolve().parent.parent / "models"


def inputs(mech):
    """(jslot, spc) index arrays of the model inputs for a mechanism."""
    return (np.array(J_SLOTS, dtype=np.int32),
            np.array([mech.species.index(s) for s in SPECIES], dtype=np.int32))


def labels(hlog, nstep, dt):
    """ln(h) label per cell from a step log hlog[2*NLOG][ncell] (float) and
    the number of attempts per cell nstep[ncell] (ISTATUS(3))."""
    nlog = hlog.shape[0] // 2
    H = hlog[0::2].astype(np.float64)
    E = hlog[1::2].astype(np.float64)
    n = hlog.shape[1]
    valid = np.arange(nlog)[:, None] < np.minimum(nstep, nlog)[None, :]
    limited = valid & (E > (FAC_SAFE / FAC_MAX) ** 3)
    first = np.where(limited.any(axis=0), limited.argmax(axis=0), -1)
    cols = np.arange(n)
    k = np.maximum(first, 0)
    fac = np.minimum(FAC_MAX, np.maximum(FAC_MIN, FAC_SAFE / np.maximum(E[k, cols], 1e-300) ** (1 / 3)))
    h_lim = H[k, cols] * fac
    last = np.maximum(np.minimum(nstep, nlog) - 1, 0)
    h_free = np.minimum(H[last, cols] * FAC_MAX, dt)
    h = np.where(first >= 0, h_lim, h_free)
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
    def raw_features(self, met, photol, conc):
        """[ncell][48] un-normalised inputs, computed by the features kernel."""
        self.net.set_norm(np.zeros(48, np.float32), np.ones(48, np.float32), 0.0, 1.0)
        X = self.net.features(met, photol, conc)
        self.net.set_norm(self.mean, self.std, self.out_mu, self.out_sigma)
        return X

    def fit_normalisation(self, Xraw, lnh, nstep):
        """Input mean/std, label mean/std and class thresholds from training data."""
        x = Xraw[:, :40].double()
        self.mean[:40] = x.mean(dim=0).cpu().numpy()
        self.std[:40] = np.maximum(x.std(dim=0).cpu().numpy(), 1e-6)
        self.out_mu, self.out_sigma = float(np.mean(lnh)), float(max(np.std(lnh), 1e-6))
        self.thresholds = [float(v) for v in np.quantile(nstep, [0.25, 0.5, 0.75])]
        self.net.set_norm(self.mean, self.std, self.out_mu, self.out_sigma)

    def normalise(self, Xraw):
        import torch
        m = torch.from_numpy(self.mean).to(Xraw.device)
        s = torch.from_numpy(self.std).to(Xraw.device)
        X = (Xraw - m) / s
        X[:, 40:] = 0.0
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
    def predict(self, met, photol, conc, safety=0.5, dt=1200.0, h0=None, cls=None):
        return self.net.predict(met, photol, conc, safety=safety, hmin=1e-5, hmax=dt, h0=h0, cls=cls)


def log_step_histories(ctx, w_dev, ic, rc, dt, nlog=NLOG):
    """Cold-start persistent step of one workload with the step log on:
    returns (hlog [2*nlog][n] float32 host, nstep [n] host, ierr)."""
    import torch
    from . import hip
    n = w_dev["conc"].shape[1]
    hlog = torch.zeros((2 * nlog, n), dtype=torch.float32, device="cuda")
    ctx.set_solver(hip.SOLVER_PERSISTENT)
    ctx.set_step_log(hlog, nlog)
    conc = w_dev["conc"].clone()
    hs = torch.zeros(n, dtype=torch.float64, device="cuda")
    hnew, ist, ierr, ok = ctx.fullchem_step(w_dev["met"], w_dev["photol"], w_dev["het"], conc, hs,
                                            dt, ic, rc)
    ctx.set_step_log(None)
    return hlog.cpu().numpy(), ist[2].cpu().numpy(), ierr.cpu().numpy()


def default_model_path(mech_name="fullchem"):
    return MODEL_DIR / ("h0_mlp_%s.npz" % mech_name)
