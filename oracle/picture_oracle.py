"""numpy restatement of the synthetic pictures: pointsPerOrientation
(datagen.jl:88-144) with getVolumeInterpolated for a SurfaceVolume
(spherical_harmonics.jl:484-530), ang_to_i (:213-221), cartesianToSpherical
(utilities.jl:115-121) and detector_to_Ewald_sphere (datagen.jl:33-39).

TEST INFRASTRUCTURE ONLY: imported by tests/ -- never by the product package.

Random numbers are the library's streams, replayed through the C oracle's
Philox4x32-10 (or_normal, or_philox4x32_10).  A picture p is made of
segments s in generateHistogram's order (datagen.jl:399-407): per particle
its photons, then, with noise photons asked for, that particle's noise
photons.  Particle n's rotation quaternion uses the normals (seed, step p,
index 0xFFFFFFFF - n, tag ROT); segment s uses normals (seed, p, s, tag PHO)
numbers 3i and 3i+1 for photon i's x and y, and one uniform (counter
{i, 0xFFFFFFFE - s, p, tag ACC}) for its acceptance; with a photon-count
variance, a particle segment's count comes from normal (seed, p, s, 0, tag
CNT).  A beamstop shot takes its two uniforms from the block {i, 0xFFFFFFFE,
p, tag BST}.  The reference draws
these from Julia's generator, so parity with it is on the acceptance rule
given the draws, not on the draws.

Pinned in tests/test_oracle_pictures.py: the interpolation returns the grid
values at grid nodes and is exactly linear in k between shells; a volume
constant in angle and radius gives photons uniform on the detector disk
(the envelope cancels); the rotations are orthonormal with det 1; the noise
volume is isotropic with c00(k) the normal density at k dr (its surface is
c00 / sqrt(4 pi) everywhere); one particle without noise is
points_per_orientation; the photon counts have the asked mean and spread;
beamstop shots fill the annulus between the beamstop and qcut uniformly.
"""
import numpy as np

from . import pyoracle as po
from .c2_oracle import surface

TAG_PROT, TAG_PPHO, TAG_PACC = 0x50524F54, 0x5050484F, 0x50414343
TAG_PCNT, TAG_PBST = 0x50434E54, 0x50425354


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


def rotation(seed, pic, part=0):
    L = po.lib()
    rl = 0xFFFFFFFF - part
    q = np.array([L.or_normal(seed, pic, rl, j, TAG_PROT) for j in range(4)])
    w, x, y, z = q / np.sqrt((q * q).sum())
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def surfaces(coeff, LMAX):
    return np.array([surface(c, LMAX).ravel() for c in coeff])


def noise_coefficients(KMAX, LMAX, rmax, sigma):
    """get_noise_volume (datagen.jl:296-304): every term deleted, c00(k) =
    gaussian_distribution(k dr; 0, sigma) (utilities.jl:129-133)."""
    c = np.zeros((KMAX, LMAX * LMAX), np.complex128)
    s2 = 2.0 * sigma ** 2
    for k in range(1, KMAX + 1):
        q = k * (rmax / KMAX)
        c[k - 1, 0] = 1 / np.sqrt(np.pi * s2) * np.exp(-q ** 2 / s2)
    return c


def _segment(surf, LMAX, KMAX, rmax, qcut, sigma, n_inc, lam, bw, seed, pic, seg, R):
    L = po.lib()
    maxI = float(np.max(surf.real)) * 1.3
    lg0 = -1.5 * np.log(2.0 * np.pi) - 3.0 * np.log(sigma)
    max_gauss = np.exp(lg0)
    acc = []
    for i in range(n_inc):
        x = sigma * L.or_normal(seed, pic, seg, 3 * i, TAG_PPHO)
        y = sigma * L.or_normal(seed, pic, seg, 3 * i + 1, TAG_PPHO)
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
        o = po.philox([i, 0xFFFFFFFE - seg, pic, TAG_PACC], [seed & 0xFFFFFFFF, seed >> 32])
        if _u01(o[0], o[1]) < ratio:
            acc.append([x, y, 0.0])
    return acc


def points_per_orientation(surf, LMAX, KMAX, rmax, qcut, sigma, n_inc, lam, bw, seed, pic):
    """(accepted (n, 3), R) of picture `pic` with one particle and no noise (datagen.jl:88-144)."""
    R = rotation(seed, pic)
    return np.array(_segment(surf, LMAX, KMAX, rmax, qcut, sigma, n_inc, lam, bw, seed, pic, 0, R)).reshape(-1, 3), R


def photon_count(seed, pic, seg, n_inc, variance):
    """pointsPerOrientation's incident_photon_variance (datagen.jl:94-99):
    round(Normal(n_inc, variance)), ties to even, at least 0."""
    if variance <= 0:
        return n_inc
    v = np.rint(n_inc + variance * po.lib().or_normal(seed, pic, seg, 0, TAG_PCNT))
    return int(v) if v > 0 else 0


def beamstop_shot(qcut, n, bw, seed, pic):
    """beamstop_shot (datagen.jl:147-161): 2 qcut (u - 1/2) per coordinate,
    kept outside the beamstop and inside |p| <= qcut."""
    acc = []
    for i in range(n):
        o = po.philox([i, 0xFFFFFFFE, pic, TAG_PBST], [seed & 0xFFFFFFFF, seed >> 32])
        x = 2 * qcut * (_u01(o[0], o[1]) - 0.5)
        y = 2 * qcut * (_u01(o[2], o[3]) - 0.5)
        if abs(x) > bw and abs(y) > bw and np.sqrt(x * x + y * y) <= qcut:
            acc.append([x, y, 0.0])
    return np.array(acc).reshape(-1, 3)


def picture(surf, nsurf, LMAX, KMAX, rmax, qcut, sigma, n_inc, lam, bw, seed, pic, particles, n_noise, noise_sigma,
            variance=0):
    """All photons of picture `pic` in generateHistogram's order (datagen.jl:399-407):
    per particle its photons (count drawn with `variance`), then (n_noise > 0)
    its noise photons from nsurf with the same rotation and envelope
    1.05 noise_sigma."""
    acc = []
    seg = 0
    for part in range(particles):
        R = rotation(seed, pic, part)
        n = photon_count(seed, pic, seg, n_inc, variance)
        acc += _segment(surf, LMAX, KMAX, rmax, qcut, sigma, n, lam, bw, seed, pic, seg, R)
        seg += 1
        if n_noise > 0:
            acc += _segment(nsurf, LMAX, KMAX, rmax, qcut, 1.05 * noise_sigma, n_noise, lam, bw, seed, pic, seg, R)
            seg += 1
    return np.array(acc).reshape(-1, 3)
