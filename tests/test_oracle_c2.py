"""The two-photon oracle (oracle/c2_oracle.py) pinned without the library.

two_photons: against the pair average <f_k1(n1) f_k2(n2)> over all pairs of
directions at angle alpha, computed by quadrature from the shells' surfaces
(Funk-Hecke makes that average equal to twoPhotons' series, scale 1/(4 pi)
included); against an l = 0 closed form; and invariance under rotations of
the l-blocks.  retrieve_solution: the round trip c2 -> structure -> c2, the
rank of G, and the negativity measure on closed-form surfaces."""
import numpy as np
import pytest
from scipy.special import sph_harm_y

from helpers import random_real_coeff, seanindex


@pytest.fixture(scope="module")
def c2o(oracle):
    from oracle import c2_oracle
    return c2_oracle


def restricted_coeff(K, L, LMAX, seed, c00=4.0):
    """A real structure with no l > 2k - 2 on shell k (what retrieveSolution can fit)."""
    c = random_real_coeff(K, L, LMAX, seed, c00=c00, decay=0.1)
    for k in range(1, K + 1):
        for l in range(2 * k - 1, LMAX):
            for m in range(-l, l + 1):
                c[k - 1, seanindex(l, m, LMAX) - 1] = 0.0
    return c


def surface_at(c, LMAX, theta, phi):
    f = np.zeros(np.shape(theta), np.complex128)
    for l in range(LMAX):
        for m in range(-l, l + 1):
            v = c[seanindex(l, m, LMAX) - 1]
            if v != 0:
                f = f + v * sph_harm_y(l, m, theta, phi)
    return f.real


def test_two_photons_is_the_pair_average(c2o):
    L, LMAX, N, K = 6, 7, 8, 3
    c = random_real_coeff(K, L, LMAX, 11, c00=2.0, decay=0.1)
    got = c2o.two_photons(c, L, LMAX, N, K)
    # exact quadrature for band-limited integrands: Gauss-Legendre in cos(theta1),
    # uniform in phi1 and in the angle psi of n2 around n1
    x, w = np.polynomial.legendre.leggauss(10)
    nphi, npsi = 16, 16
    th1 = np.arccos(x)[:, None, None]
    ph1 = (2 * np.pi * np.arange(nphi) / nphi)[None, :, None]
    psi = (2 * np.pi * np.arange(npsi) / npsi)[None, None, :]
    wt = (w[:, None, None] * np.ones((1, nphi, npsi))) / (2 * nphi * npsi)
    TH, PH = np.broadcast_arrays(th1, ph1, psi)[:2]
    e1 = np.stack([np.sin(TH) * np.cos(PH), np.sin(TH) * np.sin(PH), np.cos(TH)], -1)
    t1 = np.stack([np.cos(TH) * np.cos(PH), np.cos(TH) * np.sin(PH), -np.sin(TH)], -1)
    t2 = np.stack([-np.sin(PH), np.cos(PH), 0 * PH], -1)
    al = c2o.alpharange(N)
    f1 = {k: surface_at(c[k], LMAX, TH, PH) for k in range(K)}
    for a, alpha in enumerate(al):
        n2 = np.cos(alpha) * e1 + np.sin(alpha) * (np.cos(psi)[..., None] * t1 + np.sin(psi)[..., None] * t2)
        th2 = np.arccos(np.clip(n2[..., 2], -1, 1))
        ph2 = np.arctan2(n2[..., 1], n2[..., 0])
        f2 = {k: surface_at(c[k], LMAX, th2, ph2) for k in range(K)}
        for k1 in range(K):
            for k2 in range(K):
                avg = (wt * f1[k1] * f2[k2]).sum()
                assert got[a, k2, k1] == pytest.approx(avg, rel=1e-10, abs=1e-12 * abs(got).max())


def test_two_photons_l0_and_layout(c2o):
    L, LMAX, N, K = 4, 5, 6, 4
    c = np.zeros((K, LMAX * LMAX), complex)
    a = np.array([1.5, -0.5, 2.0, 0.25])
    c[:, 0] = a
    got = c2o.two_photons(c, L, LMAX, N, K)
    np.testing.assert_allclose(got, np.broadcast_to((a[None, :] * a[:, None] / (4 * np.pi))[None], got.shape),
                               rtol=1e-14)
    m = c2o.two_photons(random_real_coeff(K, L, LMAX, 3), L, LMAX, N, K, minimal=True)
    for k1 in range(K):
        for k2 in range(K):
            assert (np.abs(m[:, k2, k1]).max() == 0) == (k2 > k1)
    n = c2o.two_photons(c, L, LMAX, N, K, normalize=True)
    assert np.abs(n).sum() == pytest.approx(1.0, rel=1e-14)


def test_two_photons_invariant_under_block_rotations(c2o):
    from oracle.pyoracle import umat
    L, LMAX, N, K = 8, 9, 10, 5
    c = random_real_coeff(K, L, LMAX, 5)
    ref = c2o.two_photons(c, L, LMAX, N, K)
    rng = np.random.default_rng(1)
    r = c.copy()
    for l in range(2, L + 1, 2):
        q, _ = np.linalg.qr(rng.normal(size=(2 * l + 1, 2 * l + 1)))
        U = umat(l)
        ix = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        r[:, ix] = (U.conj().T @ q @ U @ c[:, ix].T).T
    np.testing.assert_allclose(c2o.two_photons(r, L, LMAX, N, K), ref, rtol=1e-11, atol=1e-13 * np.abs(ref).max())


def test_surface_closed_forms(c2o):
    LMAX = 4
    c = np.zeros(LMAX * LMAX, complex)
    c[0] = 2.0
    f = c2o.surface(c, LMAX)
    np.testing.assert_allclose(f, 2.0 / np.sqrt(4 * np.pi), rtol=1e-14)
    c[:] = 0
    c[seanindex(1, 0, LMAX) - 1] = 1.0
    th = np.pi * (2 * np.arange(2 * LMAX) + 1) / (4 * LMAX)
    np.testing.assert_allclose(c2o.surface(c, LMAX).real, np.sqrt(3 / (4 * np.pi)) * np.cos(th)[:, None] *
                               np.ones((1, 2 * LMAX)), atol=1e-14)
    full = np.zeros((2, LMAX * LMAX), complex)
    full[0, 0], full[1, 0] = 1.0, -1.0
    assert c2o.negativity(full, LMAX, 1, 1) == 0.0
    assert c2o.negativity(full, LMAX, 2, 2) == pytest.approx(1.0)


@pytest.mark.parametrize("L,LMAX,N,K2,lo,c00", [(4, 6, 8, 10, 1, 4.0), (8, 10, 16, 24, 1, 4.0),
                                                 (8, 10, 16, 24, 1, 30.0), (4, 6, 12, 14, 3, 4.0)])
def test_retrieve_round_trip(c2o, L, LMAX, N, K2, lo, c00):
    """c2 of the retrieved structure equals the c2 it was fitted to, on the
    fitted shells, for a structure the fit can represent (no l > 2k - 2 on
    shell k); G has rank <= 2l+1 and its top eigenvalues are positive.  The
    sign flip follows the summed negativity (a sum over shells, so rough
    random structures flip); a nearly flat positive structure keeps its signs."""
    c = restricted_coeff(K2, L, LMAX, 7 + L, c00)
    c2 = c2o.two_photons(c, L, LMAX, N, K2)
    rec, info = c2o.retrieve_solution(c2, L, LMAX, lo, K2)
    back = c2o.two_photons(rec, L, LMAX, N, K2)
    s = slice(lo - 1, K2)
    np.testing.assert_allclose(back[:, s, s], c2[:, s, s], rtol=1e-9, atol=1e-10 * np.abs(c2).max())
    for l in range(0, L + 1, 2):
        ev = np.linalg.eigvalsh(info["G"][l])
        top = np.sort(ev)[::-1]
        assert np.all(np.abs(top[2 * l + 1:]) <= 1e-10 * top[0])
        assert info["w"][l][-1] > 0
    assert info["flipped"] == (info["negativity"] > 0.33)
    before = -rec if info["flipped"] else rec
    assert c2o.negativity(before, LMAX, lo, K2) == pytest.approx(info["negativity"], rel=1e-12)
    if c00 >= 30.0:
        assert not info["flipped"] and info["negativity"] == 0.0
        assert np.all(rec[lo - 1:, 0].real > 0)


def test_two_photons_ewald_is_the_rotation_average(c2o):
    """lambda != 0: c2(a, k2, k1) = (1/8 pi^2) int_SO(3) f_k1(R n1) f_k2(R n2) dR
    with n1, n2 on the Ewald sphere (polar angles acos(k dq lambda / 4 pi),
    azimuths 0 and alpha_a); exact quadrature over Euler angles."""
    L, LMAX, N, K, lam, dq = 4, 5, 6, 3, 2.0, 1.1
    c = random_real_coeff(K, L, LMAX, 13, c00=2.0)
    got = c2o.two_photons(c, L, LMAX, N, K, lam=lam, dq=dq)
    flat = c2o.two_photons(c, L, LMAX, N, K)
    assert np.abs(got - flat).max() > 1e-3 * np.abs(flat).max()
    gx, gw = np.polynomial.legendre.leggauss(6)
    na = 10
    A, B, G = np.meshgrid(2 * np.pi * np.arange(na) / na, np.arccos(gx), 2 * np.pi * np.arange(na) / na, indexing="ij")
    W = np.broadcast_to(gw[None, :, None], A.shape) * (2 * np.pi / na) ** 2 / (8 * np.pi ** 2)
    ca, sa, cb, sb, cg, sg = np.cos(A), np.sin(A), np.cos(B), np.sin(B), np.cos(G), np.sin(G)
    # R = Rz(A) Ry(B) Rz(G)
    Rm = np.stack([np.stack([ca * cb * cg - sa * sg, -ca * cb * sg - sa * cg, ca * sb], -1),
                   np.stack([sa * cb * cg + ca * sg, -sa * cb * sg + ca * cg, sa * sb], -1),
                   np.stack([-sb * cg, sb * sg, cb], -1)], -2)

    def f(k, n):
        v = Rm @ n
        th, ph = np.arccos(np.clip(v[..., 2], -1, 1)), np.arctan2(v[..., 1], v[..., 0])
        return surface_at(c[k], LMAX, th, ph).real

    for a, alpha in enumerate(c2o.alpharange(N)):
        for k1 in range(1, K + 1):
            for k2 in range(1, K + 1):
                t1, t2 = np.arccos(k1 * dq * lam / (4 * np.pi)), np.arccos(k2 * dq * lam / (4 * np.pi))
                n1 = np.array([np.sin(t1), 0.0, np.cos(t1)])
                n2 = np.array([np.sin(t2) * np.cos(alpha), np.sin(t2) * np.sin(alpha), np.cos(t2)])
                avg = (W * f(k1 - 1, n1) * f(k2 - 1, n2)).sum()
                assert got[a, k2 - 1, k1 - 1] == pytest.approx(avg, rel=1e-10, abs=1e-12 * np.abs(got).max())


def test_retrieve_round_trip_ewald(c2o):
    """lambda != 0: the fit uses alpha_star columns (dq = qmax / hi) and no
    symmetrisation; c2 of the retrieved structure equals the c2 it was fitted to."""
    L, LMAX, N, K2, qmax, lam = 4, 6, 10, 12, 1.2, 3.0
    c = restricted_coeff(K2, L, LMAX, 17, 4.0)
    c2 = c2o.two_photons(c, L, LMAX, N, K2, lam=lam, dq=qmax / K2)
    rec, info = c2o.retrieve_solution(c2, L, LMAX, 1, K2, qmax=qmax, lam=lam)
    back = c2o.two_photons(rec, L, LMAX, N, K2, lam=lam, dq=qmax / K2)
    np.testing.assert_allclose(back, c2, rtol=1e-9, atol=1e-10 * np.abs(c2).max())
    flat, _ = c2o.retrieve_solution(c2, L, LMAX, 1, K2)
    assert np.abs(c2o.two_photons(flat, L, LMAX, N, K2, lam=lam, dq=qmax / K2) - c2).max() > 1e-6 * np.abs(c2).max()
