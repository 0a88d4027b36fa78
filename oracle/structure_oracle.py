"""numpy restatement of createSphericalHarmonicsStructure (structure.jl:124-134)
after loadPDB: getDensity (spherical_harmonics.jl:234-244) with
getDensity_cartesian (structure.jl:74-87) on getSurface's grid (i_to_ang,
spherical_harmonics.jl:203-210; sphericalToCartesian, utilities.jl:105-110),
forward (sphBesselTransform) and absoluteSquare (spherical_harmonics.jl:393-397).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the product package.

Atoms are rows (x, y, z, R, B).  Pinned in tests/test_oracle_structure.py:
a lone atom at the origin gives the isotropic Gaussian shell values (and only
c00 = sqrt(4 pi) times them), cut off at |p - a|^2 >= 5; a brute-force loop
over the grid index (the reference's i_to_ang) agrees with the vectorised
sampling; absoluteSquare of a band-limited shell synthesises to the squared
modulus; the intensity of a real density has no odd-l terms (Friedel).
"""
import numpy as np

from .c2_oracle import surface
from .phase_oracle import coefficients, sph_bessel


def grid_points(LMAX):
    """(phi, theta) of the (2 LMAX)^2 samples, theta-major (index = j S + i)."""
    S = 2 * LMAX
    i = np.arange(S)
    j = np.arange(S)
    phi = i / S * 2.0 * np.pi
    theta = np.pi * (2 * j + 1) / (2 * S)
    T, P = np.meshgrid(theta, phi, indexing="ij")
    return P.ravel(), T.ravel()


def density_surface(atoms, LMAX, K, rmax):
    """(K, (2 LMAX)^2) density samples; shell k at r = k rmax / K."""
    a = np.asarray(atoms, dtype=np.float64).reshape(-1, 5)
    phi, theta = grid_points(LMAX)
    amp = a[:, 4] / (a[:, 3] * np.sqrt(2.0 * np.pi)) ** 3
    w = 2.0 * a[:, 3] ** 2
    out = np.zeros((K, len(phi)), np.complex128)
    for k in range(1, K + 1):
        r = k * (rmax / K)
        p = np.stack([r * np.cos(phi) * np.sin(theta), r * np.sin(phi) * np.sin(theta), r * np.cos(theta)], 1)
        d = p[:, None, :] - a[None, :, :3]
        rsq = (d * d).sum(-1)
        out[k - 1] = np.where(rsq < 5.0, amp[None, :] * np.exp(-rsq / w[None, :]), 0.0).sum(1)
    return out


def absolute_square(coeff, LMAX):
    c = np.asarray(coeff, dtype=np.complex128)
    return np.array([coefficients(np.abs(surface(x, LMAX)).ravel() ** 2, LMAX) for x in c])


def structure(atoms, LMAX, K, rmax):
    """(density, fourier, intensity) coefficients, each (K, LMAX^2) seanindex layout."""
    dens = np.array([coefficients(s, LMAX) for s in density_surface(atoms, LMAX, K, rmax)])
    four = sph_bessel(dens, LMAX, True)
    return dens, four, absolute_square(four, LMAX)
