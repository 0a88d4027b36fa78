"""Seeded synthetic inputs for the benchmark configurations.

  small  : 200 point atoms in a ball of radius 8, KMAX=8, qmax=pi, L=8, K=8, N=16,
           histogram of 1e5 Poisson counts
  full   : 1000 point atoms in a ball of radius 26, KMAX=38, qmax=pi (rmax=KMAX),
           L=18, K=26, N=32, histogram of 1e9 Poisson counts; plain annealing with
           initial stepsize pi/4, decay 0.99998, step factor 1.01, T factor 0.1

Random numbers come from numpy's Philox bit generator keyed by the seed.
"""
import numpy as np

from . import SphericalHarmonicsVolume, seanindex

CONFIGS = {
    "small": dict(natoms=200, radius=8.0, atoms_seed=0, KMAX=8, qmax=np.pi, LMAX=25, L=8, K=8, N=16,
                  counts=1e5, hist_seed=10, start_seed=20, ntheta=48, nphi=96),
    "full": dict(natoms=1000, radius=26.0, atoms_seed=1, KMAX=38, qmax=np.pi, LMAX=25, L=18, K=26, N=32,
                 counts=1e9, hist_seed=11, start_seed=21, ntheta=96, nphi=192),
}


def rng(seed):
    return np.random.Generator(np.random.Philox(seed))


def atoms_in_ball(natoms, radius, seed):
    """Uniform points in a ball (rejection-free: direction x radius * u^(1/3))."""
    g = rng(seed)
    d = g.normal(size=(natoms, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = radius * g.uniform(size=natoms) ** (1.0 / 3.0)
    return d * r[:, None]


def umat(l):
    """Complex-to-real transform of order l (spherical_harmonics.jl:445-457)."""
    n = 2 * l + 1
    U = np.zeros((n, n), np.complex128)
    s = 1 / np.sqrt(2)
    for u in range(-l, l + 1):
        for m in range(-l, l + 1):
            v = 0j
            sg = (-1) ** abs(m)
            if m == 0 and u == 0:
                v += 1
            if u > 0 and m == u:
                v += s
            if u < 0 and m == u:
                v += 1j * sg * s
            if u < 0 and m == -u:
                v += -1j * s
            if u > 0 and m == -u:
                v += sg * s
            U[u + l, m + l] = v
    return U


def haar_rotation(n, g):
    """random_rotation (utilities.jl:165-176): sign-fixed QR of a Gaussian matrix, det > 0."""
    while True:
        Q, R = np.linalg.qr(g.normal(size=(n, n)))
        C = Q * np.sign(np.diag(R))
        if np.linalg.det(C) > 0:
            return C


def rotate_l(coeff, l, Rreal, LMAX):
    """x <- rtc R ctr x for every shell (cvec_get/cvec_set, spherical_harmonics.jl:187-199)."""
    U = umat(l)
    rot = U.conj().T @ Rreal.astype(np.complex128) @ U
    idx = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
    out = coeff.copy()
    out[:, idx] = coeff[:, idx] @ rot.T
    return out


def random_start_structure(vol, L, seed):
    """randomStartStructure (determination.jl:4-17): an independent Haar SO(2l+1) per even l >= 2."""
    g = rng(seed)
    c = vol.coeff.copy()
    for l in range(2, L + 1, 2):
        c = rotate_l(c, l, haar_rotation(2 * l + 1, g), vol.LMAX)
    return SphericalHarmonicsVolume(c, vol.LMAX, vol.KMAX, vol.rmax)


def poisson_histogram(c3raw, N, K, total, seed):
    """Photon-triplet histogram for the energy.

    counts ~ Poisson(total * p), p = max(c3,1e-30) / sum over every (bin, k-triple)
    cell; placed at [a, b, k3, k2, k1] for k1 >= k2 >= k3, then loadHistograms'
    treatment (datagen.jl:460-468): clamp every entry to >= 1e-30 and divide by
    the global sum.  Returns (c3ref, counts_per_column)."""
    R = 2 * N * N
    assert c3raw.shape[0] == R
    p = np.maximum(c3raw, 1e-30)
    p = p / p.sum()
    counts = rng(seed).poisson(total * p).astype(np.float64)
    h = np.zeros((N, 2 * N, K, K, K))
    i = 0
    for k1 in range(1, K + 1):
        for k2 in range(1, k1 + 1):
            for k3 in range(1, k2 + 1):
                h[:, :, k3 - 1, k2 - 1, k1 - 1] = counts[:, i].reshape(2 * N, N).T
                i += 1
    h = np.maximum(h, 1e-30)
    return h / h.sum(), counts


def beamstop_pictures(npic, nphot, qcut, seed, beamstop_width=0.0):
    """Sparse pictures as beamstop_shot makes them (datagen.jl:146-160): nphot
    uniform points of [-qcut, qcut]^2 (z = 0) per picture, those with |x| or |y|
    inside the beamstop and those beyond |p| = qcut dropped.  Returns (photons
    (n, 3), offsets (npic + 1,))."""
    g = rng(seed)
    pts = 2.0 * qcut * (g.uniform(size=(npic, nphot, 2)) - 0.5)
    keep = (np.abs(pts[..., 0]) > beamstop_width) & (np.abs(pts[..., 1]) > beamstop_width) & \
           (np.hypot(pts[..., 0], pts[..., 1]) <= qcut)
    counts = keep.sum(axis=1)
    ph = np.zeros((int(counts.sum()), 3))
    ph[:, :2] = pts[keep]
    off = np.zeros(npic + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return ph, off
