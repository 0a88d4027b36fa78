"""numpy restatement of the synthetic pictures: pointsPerOrientation
(datagen.jl:88-144) with getVolumeInterpolated for a SurfaceVolume
(spherical_harmonics.jl:484-530), ang_to_i (:213-221), cartesianToSpherical
(utilities.jl:115-121) and detector_to_Ewald_sphere (datagen.jl:33-39).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the product package.

Random numbers are the library's streams, replayed through the C oracle's
Philox4x32-10 (or_normal, or_philox4x32_10): picture p uses normals
(seed, step p, index 0xFFFFFFFF, tag ROT) for its rotation quaternion,
(seed, p, 0, tag PHO) numbers 3i and 3i+1 for photon i's x and y, and one
uniform (counter {i, 0xFFFFFFFE, p, tag ACC}) for its acceptance.  The
reference draws these from Julia's generator, so parity with it is on the
acceptance rule given the draws, not on the draws.

Pinned in tests/test_oracle_pictures.py: the interpolation returns the grid
values at grid nodes and is exactly linear in k between shells; a volume
constant in angle and radius gives photons uniform on the detector disk
(the envelope cancels); the rotations are orthonormal with det 1.
"""
import numpy as np

from . import pyoracle as po
from .c2_oracle import surface

TAG_PROT, TAG_PPHO, TAG_PACC = 0x50524F54, 0x5050484F, 0x50414343


def ang_to_i(phi, theta, LMAX):
    size = 2 * LMAX
    phi = np.mod(phi, 2.0 * np.pi)
    theta = np.mod(theta, np.pi)
    ip = int(np.rint(phi / (2.0 * np.pi) * size))
    it = int(np.rint(((theta * 2 * size) / np.pi - 1) / 2))
    return int(np.clip(it * size + ip + 1, 1, size * size))


def interpolate(surf, LMAX, KMAX, rmax, vec):
    """Re getVolumeInterpolated(SurfaceVolume, vec); surf (KMAX, (2 LMAX)^2)."""
    x, y, z = vec
    q = float(np.sqrt(x * x + y * y + z * z))
    p = float(np.arctan2(y, x))
    if p < 0.0:
        p = p + 2.0 * np.pi
    t = float(np.arccos(np.clip(z / q, -1.0, 1.0))) if q > 0.0 else 0.0
    size = 2 * LMAX
    dp, dt = 2.0 * np.pi / size, np.pi / size
    k = q / (rmax / KMAX)
    k0 = int(np.clip(np.floor(k), 1, KMAX))
    k1 = int(np.clip(np.ceil(k), 1, KMAX))
    p0, p1 = np.floor(p / dp) * dp, np.ceil(p / dp) * dp
    t0 = max(np.floor(((t * 2 * size) / np.pi - 1) / 2) * dt + dt / 2, 0)
    t1 = min(np.ceil(((t * 2 * size) / np.pi - 1) / 2) * dt + dt / 2, np.pi)
    a, b, c = k - k0, (p - p0) / dp, (t - t0) / dt
    s = lambda kk, pp, tt: surf[kk - 1][ang_to_i(pp, tt, LMAX) - 1]
    v = ((1 - a) * (1 - b) * (1 - c) * s(k0, p0, t0) + a * (1 - b) * (1 - c) * s(k1, p0, t0)
         + (1 - a) * b * (1 - c) * s(k0, p1, t0) + (1 - a) * (1 - b) * c * s(k0, p0, t1)
         + a * (1 - b) * c * s(k1, p0, t1) + (1 - a) * b * c * s(k0, p1, t1)
         + a * b * (1 - c) * s(k1, p1, t0) + a * b * c * s(k1, p1, t1))
    return float(np.real(v))


def _u01(lo, hi):
    return float(((int(hi) << 32) | int(lo)) >> 11) * (1.0 / 9007199254740992.0)


def rotation(seed, pic):
    L = po.lib()
    q = np.array([L.or_normal(seed, pic, 0xFFFFFFFF, j, TAG_PROT) for j in range(4)])
    w, x, y, z = q / np.sqrt((q * q).sum())
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def surfaces(coeff, LMAX):
    return np.array([surface(c, LMAX).ravel() for c in coeff])


def points_per_orientation(surf, LMAX, KMAX, rmax, qcut, sigma, n_inc, lam, bw, seed, pic):
    """(accepted (n, 3), R) of picture `pic` (datagen.jl:88-144)."""
    L = po.lib()
    maxI = float(np.max(surf.real)) * 1.3
    R = rotation(seed, pic)
    lg0 = -1.5 * np.log(2.0 * np.pi) - 3.0 * np.log(sigma)
    max_gauss = np.exp(lg0)
    acc = []
    for i in range(n_inc):
        x = sigma * L.or_normal(seed, pic, 0, 3 * i, TAG_PPHO)
        y = sigma * L.or_normal(seed, pic, 0, 3 * i + 1, TAG_PPHO)
        if not (abs(x) > bw and abs(y) > bw):
            continue
        qn = np.sqrt(x * x + y * y)
        if qn > qcut:
            continue
        th = np.arccos(qn * lam / (4 * np.pi))
        e = np.array([np.sin(th) * x, np.sin(th) * y, np.cos(th) * qn])
        r = R @ e
        I = interpolate(surf, LMAX, KMAX, rmax, r)
        gx = np.exp(lg0 - 0.5 * ((r[0] / sigma) ** 2 + (r[1] / sigma) ** 2 + (r[2] / sigma) ** 2))
        ratio = I / (maxI / max_gauss * gx)
        o = po.philox([i, 0xFFFFFFFE, pic, TAG_PACC], [seed & 0xFFFFFFFF, seed >> 32])
        if _u01(o[0], o[1]) < ratio:
            acc.append([x, y, 0.0])
    return np.array(acc).reshape(-1, 3), R
