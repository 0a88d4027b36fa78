"""Host algebra of the gradient-guided moves (tp_amd.descent): the batched
candidate construction equals the per-matrix form, keeps the coefficients of
a real function real, and the Riemannian direction is the steepest descent of
a test function on SO(2l+1) (checked against finite differences).  CPU only."""
import numpy as np
import pytest

from helpers import random_real_coeff, seanindex


class _B:
    L, LMAX, K = 8, 10, 4


def _moves():
    from tp_amd import descent
    return descent.GradientMoves(_B())


def _skews(gm, seed):
    rng = np.random.default_rng(seed)
    om = {}
    for l in gm.ls:
        A = rng.normal(size=(2 * l + 1, 2 * l + 1))
        om[l] = 0.5 * (A - A.T)
    return om


def test_candidates_equal_matrix_form():
    gm = _moves()
    c = random_real_coeff(4, 8, 10, 1)
    om = _skews(gm, 0)
    etas = [0.01, 0.1, 0.7]
    a = gm.candidates(c, om, etas)
    path = np.array(sorted({seanindex(l, m, 10) - 1 for l in range(0, 9, 2) for m in range(-l, l + 1)}))
    for s, R in enumerate(gm.rotations(om, etas)):
        for l, Rl in R.items():
            assert np.abs(Rl @ Rl.T - np.eye(2 * l + 1)).max() < 1e-13
            assert np.linalg.det(Rl) == pytest.approx(1.0, abs=1e-12)
        full = gm.apply(c, R)
        assert np.abs(a[s][:, path] - full[:, path]).max() < 1e-13
        assert np.abs(gm.moved(c, om, etas[s]) - full).max() < 1e-13
    x = gm.moved(c, om, etas[2])
    for k in range(4):
        for l in range(0, 9, 2):
            for m in range(1, l + 1):
                assert abs(x[k, seanindex(l, -m, 10) - 1] - (-1) ** m * np.conj(x[k, seanindex(l, m, 10) - 1])) < 1e-13


def test_riemannian_direction_is_steepest_descent():
    """For f(c) = sum_k Re(b_k . c_k) + |c|^2 (any smooth f), the derivative of f
    along exp(-eps Omega) (Omega from the gradient of f) is -|Omega|_F^2."""
    gm = _moves()
    rng = np.random.default_rng(3)
    c = random_real_coeff(4, 8, 10, 2)
    bvec = rng.normal(size=c.shape) + 1j * rng.normal(size=c.shape)

    def f(cc):
        return float(np.real(np.vdot(bvec, cc)) + 0.3 * np.sum(np.abs(cc) ** 4))

    # complex gradient df/dRe + i df/dIm, by central differences per entry
    g = np.zeros_like(c)
    h = 1e-6
    for idx in np.ndindex(c.shape):
        for unit in (1.0, 1j):
            d = np.zeros_like(c)
            d[idx] = unit * h
            v = (f(c + d) - f(c - d)) / (2 * h)
            g[idx] += v if unit == 1.0 else 1j * v
    om = gm.riemannian(c, g)
    n2 = gm.norm(om) ** 2
    eps = 1e-5
    fp, fm = f(gm.moved(c, om, eps)), f(gm.moved(c, om, -eps))
    assert (fp - fm) / (2 * eps) == pytest.approx(-n2, rel=1e-5)
