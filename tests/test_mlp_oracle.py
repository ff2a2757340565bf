"""CPU checks that pin the numpy MLP restatement (oracle/mlp.py) the GPU
MLP kernels are compared with: its gradient against central finite
differences of its own loss (float64), its parameter layout against the
C ABI's size, and its bf16 rounding against numpy's float32 bit layout."""
import numpy as np
import pytest


def test_layout_matches_abi_size():
    from oracle import mlp
    from gckpp_amd import hip
    assert mlp.NPARAM == hip.MLP_NPARAM == 41360
    assert mlp.NIN == 42 and mlp.F == 48


def test_gradient_matches_finite_differences():
    from oracle import mlp
    rng = np.random.default_rng(0)
    p = rng.normal(0.0, 0.15, mlp.NPARAM)
    X = rng.normal(size=(32, mlp.F))
    X[:, mlp.NIN:] = 0.0
    yh = rng.normal(size=32)
    yc = rng.integers(0, 4, 32)
    L, g = mlp.loss_grad(p, X, yh, yc, wh=0.7, wc=1.3)
    idx = np.concatenate([rng.choice(np.arange(a, b), 6, replace=False) for a, b in mlp.OFFS]
                         + [np.array([o[1] + 3 for o in mlp.OFFS])])   # weights and biases
    for i in idx:
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (mlp.loss_grad(pp, X, yh, yc, 0.7, 1.3)[0] - mlp.loss_grad(pm, X, yh, yc, 0.7, 1.3)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding():
    from oracle import mlp
    x = np.array([1.0, 1.00390625, 1.005859375, -3.14159, 65504.0, 1e-30], dtype=np.float32)
    r = mlp.bf16(x)
    # representable values stay, ties go to even, others to nearest
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert abs(r[3] - -3.140625) < 1e-12
    u = r.astype(np.float32).view(np.uint32)
    assert np.all((u & 0xFFFF) == 0)
