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


def test_expm_skew_is_the_exact_exponential():
    """expm_skew (the device's scaling-and-squaring Taylor form) against the
    exponential by eigen-decomposition, over angles from 1e-6 to 3 pi:
    agreement and orthogonality to 1e-13, det 1."""
    from tp_amd import descent
    rng = np.random.default_rng(7)
    for n in (5, 17, 37):
        A = rng.normal(size=(n, n))
        A = 0.5 * (A - A.T)
        A /= np.linalg.norm(A)
        w, V = np.linalg.eigh(1j * A)
        for th in (1e-6, 0.01, 0.3, 1.0, 2.5, 3 * np.pi):
            R = descent.expm_skew(-th * A)
            ref = ((V * np.exp(1j * th * w)) @ V.conj().T).real
            assert np.abs(R - ref).max() < 1e-13, (n, th)
            assert np.abs(R @ R.T - np.eye(n)).max() < 1e-13
            assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-12)


def test_geodesic_rotations_equal_angle_rotations():
    """The generator form (what the device receives) gives the rotations of the
    eigen form per candidate and per l; theta = 0 is the identity."""
    from tp_amd import descent
    gm = descent.GeodesicMoves(_B())
    om = _skews(gm, 4)
    gen, norms = gm.generators(om)
    assert gen.size == sum((2 * l + 1) ** 2 for l in gm.ls)
    assert np.allclose(norms, [np.linalg.norm(om[l]) for l in gm.ls])
    theta = np.random.default_rng(5).uniform(0, 2, size=(6, len(gm.ls)))
    theta[0] = 0.0
    a, b = gm.geodesic_rotations(gen, theta), gm.angle_rotations(om, theta)
    for l in gm.ls:
        assert np.abs(a[l] - b[l]).max() < 1e-13
        assert np.abs(a[l][0] - np.eye(2 * l + 1)).max() < 1e-15


def test_plane_draws_are_rotations_of_the_reference_move():
    """(u, v, b) rows: u, v orthonormal, the plane rotation orthogonal with det 1,
    turning u by b towards v and leaving the complement fixed."""
    from tp_amd import descent
    gm = descent.GeodesicMoves(_B())
    rng = np.random.default_rng(6)
    st = np.array([0.3 * (l - 1) for l in gm.ls])
    q = descent._plane_draws(len(gm.ls), gm.ls, st, rng, 5, 8)
    assert q.shape == (5, len(gm.ls), 2 * 17 + 1)
    for s in range(5):
        for li, l in enumerate(gm.ls):
            n = 2 * l + 1
            u, v, b = q[s, li, :n], q[s, li, n:2 * n], q[s, li, 2 * n]
            assert abs(u @ u - 1) < 1e-14 and abs(v @ v - 1) < 1e-14 and abs(u @ v) < 1e-14
            R = descent.plane_rotation(u, v, b)
            assert np.abs(R @ R.T - np.eye(n)).max() < 1e-14
            assert np.abs(R @ u - (np.cos(b) * u + np.sin(b) * v)).max() < 1e-14
            z = rng.normal(size=n)
            z -= (z @ u) * u + (z @ v) * v
            assert np.abs(R @ z - z).max() < 1e-14


def test_move_log_pairs_every_shell_with_every_label():
    from tp_amd import descent
    log = descent.MoveLog(10)
    f = np.arange(3 * 192, dtype=np.float32).reshape(3, 192)
    t = np.arange(2 * 36, dtype=np.float32).reshape(2, 36)
    log.add(f, t)
    X, T = log.arrays()
    assert X.shape == (6, 192) and T.shape == (6, 36)
    assert np.array_equal(X[4], f[1]) and np.array_equal(T[4], t[1])
    log.add(f, t)   # 4 rows of room left
    assert log.n == 10 and log.full()
    assert np.array_equal(log.X[9], f[0]) and np.array_equal(log.T[9], t[1])
