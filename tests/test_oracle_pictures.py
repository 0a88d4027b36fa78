"""The picture oracle (oracle/picture_oracle.py) pinned without the library:
interpolation at grid nodes and between shells, the sampling law on a
constant volume, and the rotations."""
import numpy as np
import pytest

from helpers import random_real_coeff


@pytest.fixture(scope="module")
def pic(oracle):
    from oracle import picture_oracle
    return picture_oracle


def test_interpolation_at_nodes_and_between_shells(pic):
    LMAX, K, rmax = 5, 6, 3.0
    c = random_real_coeff(K, LMAX - 1, LMAX, 4, even=False)
    surf = pic.surfaces(c, LMAX)
    S, dr = 2 * LMAX, rmax / K
    rng = np.random.default_rng(0)
    for _ in range(30):
        k = int(rng.integers(1, K + 1))
        i = int(rng.integers(0, S))
        j = int(rng.integers(0, S))
        phi, th = 2 * np.pi * i / S, np.pi * (2 * j + 1) / (2 * S)
        v = k * dr * np.array([np.cos(phi) * np.sin(th), np.sin(phi) * np.sin(th), np.cos(th)])
        got = pic.interpolate(surf, LMAX, K, rmax, v)
        assert got == pytest.approx(surf[k - 1][j * S + i].real, rel=1e-10, abs=1e-12)
    # l = 0 only, c00 linear in k: exactly linear between shells
    c0 = np.zeros((K, LMAX * LMAX), complex)
    c0[:, 0] = 2.0 + 0.5 * np.arange(1, K + 1)
    s0 = pic.surfaces(c0, LMAX)
    for kf in (1.25, 2.5, 4.75):
        v = kf * dr * np.array([0.3, -0.5, 0.81]) / np.linalg.norm([0.3, -0.5, 0.81])
        assert pic.interpolate(s0, LMAX, K, rmax, v) == pytest.approx((2.0 + 0.5 * kf) / np.sqrt(4 * np.pi), rel=1e-12)


def test_constant_volume_gives_a_uniform_disk(pic):
    """I constant: the acceptance ratio exp(|r|^2 / 2 sigma^2) / 1.3 cancels the
    incident Gaussian (it stays below 1 for sigma = 2 qcut), so accepted photons
    are uniform on the detector disk |p| <= qcut."""
    LMAX, K, qcut = 3, 5, 1.0
    c = np.zeros((K, LMAX * LMAX), complex)
    c[:, 0] = 3.0
    surf = pic.surfaces(c, LMAX)
    pts = np.vstack([pic.points_per_orientation(surf, LMAX, K, qcut, qcut, 2 * qcut, 500, 0.0, 0.0, 11, p)[0]
                     for p in range(20)])
    r = np.linalg.norm(pts[:, :2], axis=1)
    assert len(pts) > 600 and np.all(r <= qcut) and np.all(pts[:, 2] == 0)
    assert abs(np.mean(r < qcut / np.sqrt(2)) - 0.5) < 0.06
    ang = np.arctan2(pts[:, 1], pts[:, 0])
    assert abs(np.mean(np.cos(ang))) < 0.08 and abs(np.mean(np.sin(ang))) < 0.08


def test_rotations(pic):
    for p in range(5):
        R = pic.rotation(7, p)
        np.testing.assert_allclose(R @ R.T, np.eye(3), atol=1e-14)
        assert np.linalg.det(R) == pytest.approx(1.0, abs=1e-14)


def test_noise_volume(pic):
    """get_noise_volume: only c00, the normal density of width sigma at k dr,
    so the surfaces are isotropic with value c00 / sqrt(4 pi); the density
    integrates to 1 over the real line."""
    from scipy.stats import norm
    LMAX, K, rmax, sigma = 4, 9, 2.0, 0.6
    c = pic.noise_coefficients(K, LMAX, rmax, sigma)
    assert np.count_nonzero(c[:, 1:]) == 0
    k = np.arange(1, K + 1)
    assert np.allclose(c[:, 0].real, norm.pdf(k * rmax / K, scale=sigma), rtol=1e-13)
    surf = pic.surfaces(c, LMAX)
    assert np.allclose(surf, c[:, :1] / np.sqrt(4 * np.pi), rtol=1e-12, atol=1e-15)


def test_single_particle_picture_is_points_per_orientation(pic):
    LMAX, K, rmax = 4, 6, 1.0
    c = random_real_coeff(K, LMAX - 1, LMAX, 8, even=False)
    surf = pic.surfaces(c, LMAX)
    a, _ = pic.points_per_orientation(surf, LMAX, K, rmax, rmax, rmax / 3, 60, 0.0, 0.0, 3, 2)
    b = pic.picture(surf, None, LMAX, K, rmax, rmax, rmax / 3, 60, 0.0, 0.0, 3, 2, 1, 0, 1.0)
    assert len(a) > 0 and np.array_equal(a, b)


def test_photon_count_law(pic):
    """round(Normal(n, v)): mean n, standard deviation v, integers."""
    c = np.array([pic.photon_count(4, p, 0, 200, 15) for p in range(4000)])
    assert abs(c.mean() - 200) < 4 * 15 / np.sqrt(len(c))
    assert abs(c.std() - 15) < 0.6
    assert pic.photon_count(4, 3, 0, 200, 0) == 200


def test_beamstop_shot_is_uniform_on_the_annulus(pic):
    """Uniform on the square then filtered: the kept fraction is the area of
    the disk of radius qcut outside the beamstop cross over (2 qcut)^2, and the
    radii follow r dr."""
    qcut, bw, n = 2.0, 0.3, 20000
    ph = pic.beamstop_shot(qcut, n, bw, 11, 0)
    assert np.all(np.abs(ph[:, :2]) > bw) and np.all(np.hypot(ph[:, 0], ph[:, 1]) <= qcut)
    # area of the disk minus the cross |x| <= bw or |y| <= bw (inside the disk)
    g = np.linspace(-qcut, qcut, 2001)
    X, Y = np.meshgrid(g, g)
    inside = (np.hypot(X, Y) <= qcut) & (np.abs(X) > bw) & (np.abs(Y) > bw)
    frac = inside.mean()
    assert abs(len(ph) / n - frac) < 4 * np.sqrt(frac * (1 - frac) / n) + 2e-3
    # without a beamstop, P(r < qcut / 2) = 1/4
    ph0 = pic.beamstop_shot(qcut, n, 0.0, 12, 0)
    r = np.hypot(ph0[:, 0], ph0[:, 1])
    assert abs((r < qcut / 2).mean() - 0.25) < 4 * np.sqrt(0.25 * 0.75 / len(r))
