"""The phase-retrieval oracle (oracle/phase_oracle.py) pinned without the
library: quadrature weights, the SH transform pair, the continuous limits of
the spherical Bessel transforms, and sRAAR's constraints."""
import numpy as np
import pytest
from scipy.integrate import quad
from scipy.special import spherical_jn

from helpers import random_real_coeff, seanindex


@pytest.fixture(scope="module")
def po(oracle):
    from oracle import phase_oracle
    return phase_oracle


@pytest.mark.parametrize("LMAX", [3, 8, 25])
def test_dh_weights(po, LMAX):
    w = po.dh_weights(LMAX)
    th = po.thetas(LMAX)
    x = np.cos(th)
    for n in range(2 * LMAX):
        assert (w * x ** n).sum() == pytest.approx((1 - (-1) ** (n + 1)) / (n + 1), abs=1e-12)
    closed = (2.0 / LMAX) * np.sin(th) * sum(np.sin((2 * k + 1) * th) / (2 * k + 1) for k in range(LMAX))
    np.testing.assert_allclose(w, closed, rtol=0, atol=1e-14)


def test_transform_round_trip(po):
    LMAX, K = 7, 3
    c = random_real_coeff(K, LMAX - 1, LMAX, 3, even=False)
    c[1] += 0.3j * random_real_coeff(1, LMAX - 1, LMAX, 4, even=False)[0]   # not a real function
    for k in range(K):
        f = po.surface(c[k], LMAX)
        assert f.shape == (2 * LMAX, 2 * LMAX)
        np.testing.assert_allclose(po.coefficients(f.ravel(), LMAX), c[k], rtol=0, atol=1e-13 * np.abs(c[k]).max())


@pytest.mark.parametrize("l", [0, 2])
def test_bessel_continuous_limits(po, l):
    """forward(f)[k] -> i^l 8 pi / sqrt(K) int r^2 j_l(q_k r) f(r) dr and
    backward(g)[r] -> (-i)^l 2 / (pi sqrt(K)) int q^2 j_l(q r) g(q) dq for
    smooth profiles that vanish at the ends of the grids."""
    K, LMAX = 96, 3
    r = np.arange(1.0, K + 1)
    q = np.pi / K * r
    sig = 9.0
    prof = lambda x: x ** (2 + l) * np.exp(-(x / sig) ** 2)
    c = np.zeros((K, LMAX * LMAX), complex)
    s = seanindex(l, 0, LMAX) - 1
    c[:, s] = prof(r)
    fw = po.sph_bessel(c, LMAX, True)[:, s]
    for k in (4, 10, 20):
        ref = quad(lambda x: x * x * spherical_jn(l, q[k - 1] * x) * prof(x), 0, 12 * sig, limit=400)[0]
        want = (1j) ** l * 8 * np.pi / np.sqrt(K) * ref
        assert abs(fw[k - 1] - want) <= 2e-3 * np.abs(fw).max()
    sq = 0.6
    gprof = lambda x: x ** (2 + l) * np.exp(-(x / sq) ** 2)
    c[:, s] = gprof(q)
    bw = po.sph_bessel(c, LMAX, False)[:, s]
    for rr in (2, 5, 9):
        ref = quad(lambda x: x * x * spherical_jn(l, x * rr) * gprof(x), 0, 12 * sq, limit=400)[0]
        want = (-1j) ** l * 2 / (np.pi * np.sqrt(K)) * ref
        assert abs(bw[rr - 1] - want) <= 2e-3 * np.abs(bw).max()


def test_sraar_constraints(po):
    LMAX, K = 5, 6
    I = random_real_coeff(K, LMAX - 1, LMAX, 5, even=True, c00=5.0)
    ph = np.random.default_rng(0).uniform(0, 2 * np.pi, (K, 4 * LMAX * LMAX))
    out, d = po.sraar(I, LMAX, 4, 0.75, 0.75, 100.0, 0.5, ph)
    Koff = 3
    assert np.all(out[Koff:] == 0)
    assert np.all(out[:Koff, 0].real > 0)          # c00 of a nonnegative surface
    assert len(d) == 4 and all(np.isfinite(d))
    out0, d0 = po.sraar(I, LMAX, 0, 0.75, 0.75, 100.0, 0.5, ph)
    assert d0 == [] and np.all(out0[Koff:] == 0)


def _sraar_start(po, I, LMAX, ph):
    amp = np.array([np.sqrt(np.maximum(po.surface(I[k], LMAX).real.ravel(), 0.0)) for k in range(I.shape[0])])
    x0 = po._backward_surf(amp * np.exp(1j * ph), LMAX)
    return amp, x0


def test_sraar_one_step_at_beta_0_and_1(po):
    """RAAR x+ = beta/2 (R_S R_M + I) x + (1 - beta) P_M x (phases.jl:68-70 with
    R_M = 2B - I, P_S = A): one step at beta = 0 is the modulus step P_M x0 =
    B(x0) whatever the support projection does, and at beta = 1 it is
    x0 + A(R_M x0) - B(x0).  The first real-space change pins both (a loop with
    beta and 1 - beta exchanged, a wrong reflection or a dropped term fails)."""
    LMAX, K = 5, 6
    I = random_real_coeff(K, LMAX - 1, LMAX, 9, even=True, c00=5.0, decay=0.1)
    ph = np.random.default_rng(4).uniform(0, 2 * np.pi, (K, 4 * LMAX * LMAX))
    amp, x0 = _sraar_start(po, I, LMAX, ph)
    Bx = po._B(x0, amp, LMAX)
    _, d0 = po.sraar(I, LMAX, 1, 0.0, 0.0, 100.0, 0.5, ph)
    want0 = np.abs(x0 - Bx).sum() / np.abs(x0).sum()
    assert d0[0] == pytest.approx(want0, rel=1e-12)
    x1 = x0 + po._A(2.0 * Bx - x0, 3) - Bx
    _, d1 = po.sraar(I, LMAX, 1, 1.0, 1.0, 100.0, 0.5, ph)
    want1 = np.abs(x0 - x1).sum() / np.abs(x0).sum()
    assert d1[0] == pytest.approx(want1, rel=1e-12)
    assert abs(want0 - want1) > 1e-3 * want0      # the support projection is active: the two differ


def test_sraar_schedule_limits(po):
    """beta(n) = e beta0 + (1 - e) beta_max with e = exp(-(n/tau)^3): tau -> 0
    runs beta_max from the first iteration, tau -> infinity keeps beta0."""
    LMAX, K, n = 5, 6, 5
    I = random_real_coeff(K, LMAX - 1, LMAX, 11, even=True, c00=5.0, decay=0.1)
    ph = np.random.default_rng(6).uniform(0, 2 * np.pi, (K, 4 * LMAX * LMAX))
    fast = po.sraar(I, LMAX, n, 0.2, 0.9, 1e-9, 0.5, ph)
    slow = po.sraar(I, LMAX, n, 0.2, 0.9, 1e12, 0.5, ph)
    at_max = po.sraar(I, LMAX, n, 0.9, 0.9, 100.0, 0.5, ph)
    at_0 = po.sraar(I, LMAX, n, 0.2, 0.2, 100.0, 0.5, ph)
    assert np.array_equal(fast[0], at_max[0]) and fast[1] == at_max[1]
    assert np.array_equal(slow[0], at_0[0]) and slow[1] == at_0[1]
    assert not np.allclose(at_max[0], at_0[0])
