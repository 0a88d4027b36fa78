"""The oracle's proposer-MLP backward pass, pinned by finite differences.

or_mlp_loss_grad (FP64, no bf16 emulation) must match central differences of
its own loss; with emulation on it must stay close to the exact gradient (the
bf16 rounding points change it by ~1e-2 relative at most)."""
import numpy as np
import pytest


def _params(po, rng, scale=1.0):
    n = po.mlp_nparam()
    p = np.zeros(n, np.float32)
    off = 0
    for rows, cols in [(256, 192), (256, 256), (256, 256), (64, 256)]:
        w = rng.normal(size=(rows, cols)) * scale * np.sqrt(2.0 / cols)
        if rows == 64:
            w[36:] = 0.0
        p[off:off + rows * cols] = w.ravel()
        off += rows * cols
        b = rng.normal(size=rows) * 0.1
        if rows == 64:
            b[36:] = 0.0
        p[off:off + rows] = b
        off += rows
    assert off == n
    return p


def test_mlp_grad_finite_differences(oracle):
    po = oracle
    rng = np.random.default_rng(0)
    p = _params(po, rng)
    B = 3
    X = rng.normal(size=(B, 192)).astype(np.float32)
    X[:, 191] = 0.0
    T = rng.normal(size=(B, 36)).astype(np.float32)
    loss, g, _ = po.mlp_loss_grad(p, X, T, emulate=False)
    assert loss > 0
    n = p.size
    # 12 indices in each of the 8 parameter blocks
    bounds = np.cumsum([0, 256 * 192, 256, 65536, 256, 65536, 256, 64 * 256, 64])
    idx = []
    for a, b in zip(bounds[:-1], bounds[1:]):
        idx += list(rng.integers(a, b, size=12))
    bad = 0
    for i in idx:
        h = 1e-4   # float32 parameters: the step actually taken is pp[i] - pm[i]
        pp, pm = p.copy(), p.copy()
        pp[i] += h
        pm[i] -= h
        fd = (po.mlp_loss_grad(pp, X, T, emulate=False, want_grad=False)[0] -
              po.mlp_loss_grad(pm, X, T, emulate=False, want_grad=False)[0]) / (float(pp[i]) - float(pm[i]))
        if abs(fd - g[i]) > 1e-6 + 1e-5 * abs(g[i]):
            bad += 1
    # a ReLU kink inside [p-h, p+h] can spoil a difference quotient; allow one
    assert bad <= 1, bad
    # padding outputs carry no gradient
    w4 = g[bounds[6]:bounds[7]].reshape(64, 256)
    assert np.all(w4[36:] == 0)


def test_mlp_grad_emulated_close(oracle):
    po = oracle
    rng = np.random.default_rng(1)
    p = _params(po, rng)
    X = rng.normal(size=(8, 192)).astype(np.float32)
    T = rng.normal(size=(8, 36)).astype(np.float32)
    l0, g0, _ = po.mlp_loss_grad(p, X, T, emulate=False)
    l1, g1, _ = po.mlp_loss_grad(p, X, T, emulate=True)
    assert abs(l1 - l0) < 2e-2 * l0
    assert np.linalg.norm(g1 - g0) < 3e-2 * np.linalg.norm(g0)
