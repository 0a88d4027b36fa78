"""The structure oracle (oracle/structure_oracle.py) pinned without the
library: a lone Gaussian atom, the cutoff, a brute-force grid loop, the
squared modulus of band-limited shells and Friedel's law on crambin."""
import os

import numpy as np
import pytest

from helpers import random_real_coeff

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def so(oracle):
    from oracle import structure_oracle
    return structure_oracle


def test_lone_atom_at_the_origin(so, oracle):
    """rho(r) = B / (R sqrt(2 pi))^3 exp(-r^2 / (2 R^2)) on every sample of
    shells with r^2 < 5, 0 beyond; coefficients: only c00 = sqrt(4 pi) rho."""
    from oracle.phase_oracle import coefficients
    LMAX, K, rmax, R, B = 6, 8, 4.0, 0.7, 6.0
    s = so.density_surface([[0, 0, 0, R, B]], LMAX, K, rmax)
    for k in range(1, K + 1):
        r = k * rmax / K
        want = B / (R * np.sqrt(2 * np.pi)) ** 3 * np.exp(-r * r / (2 * R * R)) if r * r < 5 else 0.0
        assert np.allclose(s[k - 1], want, rtol=1e-14, atol=0)
        c = coefficients(s[k - 1], LMAX)
        assert abs(c[0] - np.sqrt(4 * np.pi) * want) <= 1e-13 * max(want, 1e-300) + 1e-300
        assert np.abs(c[1:]).max() <= 1e-13 * max(want, 1e-300) + 1e-300


def test_against_a_brute_force_grid_loop(so):
    """The reference's getSurface order: sample index i -> (phi, theta) by
    i_to_ang (mod / floor of the 1-based index)."""
    LMAX, K, rmax = 3, 3, 3.0
    rng = np.random.default_rng(2)
    atoms = np.column_stack([rng.uniform(-2, 2, (6, 3)), rng.choice([0.3, 0.6, 0.7, 1.0], 6),
                             rng.choice([1.0, 6.0, 8.0, 16.0], 6)])
    got = so.density_surface(atoms, LMAX, K, rmax)
    size = 2 * LMAX
    for k in range(1, K + 1):
        r = k * rmax / K
        for idx in range(1, size * size + 1):
            iphi, itheta = (idx - 1) % size, (idx - 1) // size
            phi, theta = iphi / size * 2.0 * np.pi, np.pi * (2 * itheta + 1) / (2 * size)
            p = np.array([r * np.cos(phi) * np.sin(theta), r * np.sin(phi) * np.sin(theta), r * np.cos(theta)])
            rho = 0.0
            for x, y, z, R, B in atoms:
                rsq = float(np.sum((p - [x, y, z]) ** 2))
                if rsq < 5:
                    rho += B / ((R * np.sqrt(2.0 * np.pi)) ** 3) * np.exp(-rsq / (2.0 * R ** 2))
            assert got[k - 1, idx - 1] == pytest.approx(rho, rel=1e-13, abs=1e-300)


def test_absolute_square_of_band_limited_shells(so, oracle):
    """For shells with l < LMAX / 2, |f|^2 stays below LMAX and the
    quadrature is exact: the result synthesises to |f|^2 point by point."""
    from oracle.c2_oracle import surface
    LMAX, K = 8, 3
    c = random_real_coeff(K, 3, LMAX, 5, even=False)
    got = so.absolute_square(c, LMAX)
    for k in range(K):
        f = surface(c[k], LMAX)
        assert np.allclose(surface(got[k], LMAX), np.abs(f) ** 2, rtol=0, atol=1e-12 * np.abs(f).max() ** 2)
    # a constant shell c00: |c00|^2 / (4 pi) everywhere, so c00' = |c00|^2 / sqrt(4 pi)
    c0 = np.zeros((1, LMAX * LMAX), np.complex128)
    c0[0, 0] = 2.0 - 1.0j
    g = so.absolute_square(c0, LMAX)
    assert g[0, 0] == pytest.approx(5.0 / np.sqrt(4 * np.pi), rel=1e-13)
    assert np.abs(g[0, 1:]).max() < 1e-13


def test_crambin_friedel_and_reality(so, oracle):
    """Crambin (the reference's structure test input) at a small LMAX: the
    density is real, c_{l,-m} = (-1)^m conj(c_lm); the intensity of a real
    density is centrosymmetric, so every odd-l coefficient vanishes."""
    from oracle.pyoracle import seanindex
    atoms = _crambin_atoms()
    assert len(atoms) == 670
    LMAX, K = 8, 6
    d, f, i = so.structure(atoms, LMAX, K, 12.0)
    scale = np.abs(d).max()
    for l in range(LMAX):
        for m in range(1, l + 1):
            a, b = d[:, seanindex(l, m, LMAX) - 1], d[:, seanindex(l, -m, LMAX) - 1]
            assert np.allclose(b, (-1) ** m * np.conj(a), atol=1e-12 * scale)
    odd = [seanindex(l, m, LMAX) - 1 for l in range(1, LMAX, 2) for m in range(-l, l + 1)]
    assert np.abs(i[:, odd]).max() <= 1e-11 * np.abs(i).max()
    assert np.abs(i[:, 0].imag).max() <= 1e-11 * np.abs(i).max() and np.all(i[:, 0].real > 0)


def _crambin_atoms():
    """(x, y, z, R, B) of crambin's atoms, read and centred as loadPDB does
    (columns 31-54, element 78; unweighted mean), independently of the package."""
    pos, el = [], []
    with open(os.path.join(GOLDEN, "crambin_atoms.pdb")) as f:
        for ln in f:
            pos.append([float(ln[30:38]), float(ln[38:46]), float(ln[46:54])])
            el.append(ln[77])
    pos = np.array(pos)
    pos -= pos.mean(0)
    RB = {"C": (0.7, 6.0), "N": (0.65, 7.0), "S": (1.0, 16.0), "O": (0.6, 8.0), "H": (0.3, 1.0)}
    return np.column_stack([pos, [RB[e][0] for e in el], [RB[e][1] for e in el]])
