"""Rotational fit of a structure to a reference (data_processing.jl:26-79) over
the device scorer tp_fit_*: every candidate rotation's similarity
(spherical_harmonics.jl:426-435) is computed on the GPU, many per call.

  rotateStructure(volume, theta, phi, gamma, K, L, basis)   spherical_harmonics.jl:598-612
  StructureFit(basis, reference, K_range).similarity(euler)  similarity of rotated copies
  fitStructures_full(volume, reference, num, K_range, L, basis, ...)   data_processing.jl:26-46
  fitStructures_random(volume, reference, K_range, L, stepsizefactor, repetitions, basis)
                                                             data_processing.jl:53-75
"""
import ctypes as C

import numpy as np

from . import SphericalHarmonicsVolume, check, lib, rotate
from ._lib import track


def euler_quat(theta, phi, gamma):
    """Unit quaternion (w, x, y, z) of R = Rz(phi) Ry(-theta) Rz(gamma), the
    rotation rotationMatrix(theta, phi, gamma, l) (spherical_harmonics.jl:589-595)
    applies: the rotated structure is f(R^-1 r)."""
    def mul(u, v):
        return np.array([u[0] * v[0] - u[1] * v[1] - u[2] * v[2] - u[3] * v[3],
                         u[0] * v[1] + u[1] * v[0] + u[2] * v[3] - u[3] * v[2],
                         u[0] * v[2] - u[1] * v[3] + u[2] * v[0] + u[3] * v[1],
                         u[0] * v[3] + u[1] * v[2] - u[2] * v[1] + u[3] * v[0]])
    a = np.array([np.cos(phi / 2), 0.0, 0.0, np.sin(phi / 2)])
    b = np.array([np.cos(-theta / 2), 0.0, np.sin(-theta / 2), 0.0])
    c = np.array([np.cos(gamma / 2), 0.0, 0.0, np.sin(gamma / 2)])
    return mul(mul(a, b), c)


def _coeff(v):
    return v.coeff if isinstance(v, SphericalHarmonicsVolume) else np.asarray(v, dtype=np.complex128)


def rotateStructure(volume, theta, phi, gamma, K, L, basis):
    """rotateStructure(volume, theta, phi, gamma, K, 2:2:L): the even l-blocks
    2 <= l <= L of shells 1..K rotated (FP64 Wigner-D on the device, tp_rotate);
    every other entry is kept.  basis supplies LMAX and the l range (its L must be L)."""
    if basis.L != L:
        raise ValueError("basis.L must equal L")
    c = np.array(_coeff(volume), dtype=np.complex128, copy=True)
    c[:K] = rotate(c, basis, euler_quat(theta, phi, gamma), nshell=K)
    if isinstance(volume, SphericalHarmonicsVolume):
        return SphericalHarmonicsVolume(c, volume.LMAX, volume.KMAX, volume.rmax)
    return c


class StructureFit:
    """Similarity of a fixed reference with rotated copies of a structure
    (tp_fit_*).  K_range: 1-based shells (a range, or an int K for 1..K);
    rotations act on shells 1..max(K_range) and even l in 2..basis.L."""

    def __init__(self, basis, reference, K_range):
        rng = range(1, K_range + 1) if isinstance(K_range, int) else K_range
        self.basis, self.klo, self.khi = basis, min(rng), max(rng)
        ref = np.ascontiguousarray(_coeff(reference), dtype=np.complex128)
        if ref.shape[0] < self.khi or ref.shape[1] < basis.LMAX ** 2:
            raise ValueError("reference needs max(K_range) shells of LMAX^2 coefficients")
        h = C.c_void_p()
        check(lib().tp_fit_create(basis.ctx.h, basis.h, ref.ctypes.data, ref.shape[1], self.klo, self.khi,
                                  C.byref(h)), "tp_fit_create")
        self.h = h
        track(self, 0)

    def _release(self):
        if getattr(self, "h", None):
            lib().tp_fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def set_volume(self, volume):
        c = np.ascontiguousarray(_coeff(volume), dtype=np.complex128)
        if c.shape[0] < self.khi or c.shape[1] < self.basis.LMAX ** 2:
            raise ValueError("volume needs max(K_range) shells of LMAX^2 coefficients")
        check(lib().tp_fit_set_volume(self.h, c.ctypes.data, c.shape[1]), "tp_fit_set_volume")
        return self

    def similarity(self, euler):
        """euler: (n, 3) rows (theta, phi, gamma) -> (n,) similarities."""
        e = np.ascontiguousarray(np.asarray(euler, dtype=np.float64).reshape(-1, 3))
        out = np.zeros(e.shape[0])
        check(lib().tp_fit_similarity(self.h, e.shape[0], e.ctypes.data, out.ctypes.data), "tp_fit_similarity")
        return out


def fitStructures_full(volume, reference, num, K_range, L, basis, ta=0.0, tb=np.pi / 2, pa=0.0, pb=np.pi,
                       ga=0.0, gb=np.pi):
    """Grid search over (num+1)^3 Euler triples (theta outer, gamma inner, as the
    reference's comprehension), all scored on the device.  Returns (rotated
    volume, best similarity, theta, phi, gamma); ties go to the first triple."""
    th = np.linspace(ta, tb, num + 1)
    ph = np.linspace(pa, pb, num + 1)
    gm = np.linspace(ga, gb, num + 1)
    T, P, G = np.meshgrid(th, ph, gm, indexing="ij")
    euler = np.stack([T.ravel(), P.ravel(), G.ravel()], axis=1)
    fit = StructureFit(basis, reference, K_range).set_volume(volume)
    sc = fit.similarity(euler)
    i = int(np.argmax(sc))
    t, p, g = euler[i]
    return rotateStructure(volume, t, p, g, fit.khi, L, basis), float(sc[i]), float(t), float(p), float(g)


def random_search(score, repetitions, stepsizefactor, rng):
    """The search loop of fitStructures_random (data_processing.jl:53-75) over a
    scorer score((n, 3) Euler rows) -> (n,) similarities: each repetition
    starts at phi = 2 pi u, theta = acos(u), gamma = 2 pi u and, while the step
    size (from pi, times stepsizefactor per step) exceeds pi/180, moves to
    theta + N(0,1) s/2, phi + N(0,1) s, gamma + N(0,1) s when that raises its
    similarity (from 0).  The repetitions advance together, one scorer call per
    step; each has its own step-size schedule (the reference's worker processes
    each start from pi).  Returns (best similarities, theta, phi, gamma), one
    entry per repetition."""
    if not 0.0 < stepsizefactor < 1.0:
        raise ValueError("stepsizefactor must be in (0, 1)")
    R = repetitions
    phi, theta, gamma = 2 * np.pi * rng.random(R), np.arccos(rng.random(R)), 2 * np.pi * rng.random(R)
    best = np.zeros(R)
    step = np.pi
    while step > np.pi / 180.0:
        z = rng.standard_normal((R, 3))
        nt, nph, ng = theta + z[:, 0] * step / 2, phi + z[:, 1] * step, gamma + z[:, 2] * step
        sc = score(np.stack([nt, nph, ng], axis=1))
        up = sc > best
        best = np.where(up, sc, best)
        theta, phi, gamma = np.where(up, nt, theta), np.where(up, nph, phi), np.where(up, ng, gamma)
        step *= stepsizefactor
    return best, theta, phi, gamma


def fitStructures_random(volume, reference, K_range, L, stepsizefactor, repetitions=4, basis=None, seed=0):
    """fitStructures_random (data_processing.jl:53-75) with every step's
    candidates scored on the device (random_search over StructureFit).
    Returns (rotated volume, best similarity, theta, phi, gamma)."""
    if basis is None:
        raise ValueError("basis is required")
    fit = StructureFit(basis, reference, K_range).set_volume(volume)
    best, theta, phi, gamma = random_search(fit.similarity, repetitions, stepsizefactor,
                                            np.random.default_rng(seed))
    i = int(np.argmax(best))
    return (rotateStructure(volume, theta[i], phi[i], gamma[i], fit.khi, L, basis), float(best[i]),
            float(theta[i]), float(phi[i]), float(gamma[i]))
