"""numpy (float64) restatement of the step-size MLP of
geos-chem_amd/csrc/mlp.hip: input features, forward pass, loss and its
gradient, and a bf16 emulation of the MFMA inference kernel.

Test infrastructure: only tests/ use it, as the checker of the GPU kernels
(it is pinned itself by central finite differences in tests).  The MLP has
no counterpart in the reference, which starts each cell from RCNTRL(3)
(GeosCore/fullchem_mod.F90:1016-1020).
"""
import numpy as np
from scipy.special import erf

F, H, O = 48, 128, 16
NJ, NS = 12, 24
NIN = 4 + NJ + NS + 2
# parameter layout: W[in][out] row-major then bias, layer after layer
SHAPES = [(F, H), (H, H), (H, H), (H, O)]
OFFS = []
_o = 0
for _a, _b in SHAPES:
    OFFS.append((_o, _o + _a * _b))
    _o += _a * _b + _b
NPARAM = _o

# met record rows (csrc/cell_fields.h)
MET_TEMP, MET_PRESS, MET_NUMDEN, MET_SUNCOS = 0, 1, 2, 4


def unpack(p):
    """[(W, b)] views of the flat parameter vector."""
    out = []
    for (a, b), (w0, b0) in zip(SHAPES, OFFS):
        out.append((p[w0:w0 + a * b].reshape(a, b), p[b0:b0 + b]))
    return out


def gelu(z):
    return 0.5 * z * (1.0 + erf(z / np.sqrt(2.0)))


def gelu_grad(z):
    return 0.5 * (1.0 + erf(z / np.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / np.sqrt(2.0 * np.pi)


def raw_features(jslot, spc, met, photol, conc, hprev=None):
    """[ncell][NIN] un-normalised inputs (layouts [field][ncell]); the last
    two: ln(hprev) and 1 where hprev > 0 (a previous step size), else 0."""
    n = conc.shape[1]
    x = np.zeros((n, NIN))
    x[:, 0] = met[MET_TEMP]
    x[:, 1] = np.log10(np.maximum(met[MET_PRESS], 1e-30))
    x[:, 2] = np.log10(np.maximum(met[MET_NUMDEN], 1e-30))
    x[:, 3] = met[MET_SUNCOS]
    for k, s in enumerate(jslot):
        x[:, 4 + k] = np.log10(np.maximum(photol[s], 0.0) + 1e-20)
    for k, s in enumerate(spc):
        x[:, 4 + NJ + k] = np.log10(np.maximum(conc[s], 0.0) + 1.0)
    if hprev is not None:
        warm = hprev > 0.0
        x[:, NIN - 2] = np.where(warm, np.log(np.where(warm, hprev, 1.0)), 0.0)
        x[:, NIN - 1] = warm
    return x


def features(jslot, spc, mean, std, met, photol, conc, hprev=None):
    """[ncell][F] normalised inputs, padded with zeros."""
    x = raw_features(jslot, spc, met, photol, conc, hprev)
    X = np.zeros((x.shape[0], F))
    X[:, :NIN] = (x - mean[:NIN]) * (1.0 / std[:NIN])
    return X


def forward(p, X):
    """Pre-activations and activations of every layer, and the outputs."""
    L = unpack(np.asarray(p, dtype=np.float64))
    zs, acts = [], [np.asarray(X, dtype=np.float64)]
    a = acts[0]
    for k, (W, b) in enumerate(L):
        z = a @ W + b
        if k < 3:
            zs.append(z)
            a = gelu(z)
            acts.append(a)
        else:
            return zs, acts, z


def loss_grad(p, X, yh, ycls, wh=1.0, wc=1.0):
    """Loss wh*mean((o0-yh)^2) + wc*mean(CE(o1..o4, ycls)) and its gradient
    (flat, parameter layout)."""
    zs, acts, o = forward(p, X)
    n = X.shape[0]
    e = o[:, 0] - yh
    lg = o[:, 1:5]
    mx = lg.max(axis=1, keepdims=True)
    ex = np.exp(lg - mx)
    s = ex.sum(axis=1, keepdims=True)
    ce = (np.log(s[:, 0]) + mx[:, 0] - lg[np.arange(n), ycls])
    L = wh * np.mean(e * e) + wc * np.mean(ce)
    d = np.zeros_like(o)
    d[:, 0] = 2.0 * wh * e / n
    sm = ex / s
    sm[np.arange(n), ycls] -= 1.0
    d[:, 1:5] = wc * sm / n
    g = np.zeros(NPARAM)
    Ls = unpack(np.asarray(p, dtype=np.float64))
    G = unpack(g)
    for k in range(3, -1, -1):
        W, _ = Ls[k]
        Gw, Gb = G[k]
        Gw[:] = acts[k].T @ d
        Gb[:] = d.sum(axis=0)
        if k > 0:
            d = (d @ W.T) * gelu_grad(zs[k - 1])
    return L, g


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    u = np.asarray(x, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def forward_bf16(p, X):
    """The inference kernel's arithmetic: bf16 inputs, weights and hidden
    activations, fp32-exact products summed in float64 (the kernel sums in
    fp32), biases and GELU in fp32 precision or better."""
    L = unpack(np.asarray(p, dtype=np.float32))
    a = bf16(X)
    for k, (W, b) in enumerate(L):
        z = a @ bf16(W) + b.astype(np.float64)
        if k < 3:
            a = bf16(gelu(z).astype(np.float32))
        else:
            return z
