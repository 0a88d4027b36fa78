"""Phase retrieval in the spherical-harmonics representation on the device
(tp_sraar; sRAAR, phases.jl:31-95) and the transforms it uses:

  getSurfaceVolume(volume)                      spherical_harmonics.jl:151-154 (tp_sh_surface)
  getSphericalHarmonicsVolume(surf, ...)        spherical_harmonics.jl:157-160 (tp_sh_coeff)
  forward(volume) / backward(volume)            spherical_harmonics.jl:362-377 (tp_sph_bessel)
  sRAAR(intensity, iterations, ...)             phases.jl:31-95

Surfaces are (KMAX, (2 LMAX)^2) complex arrays, samples theta-major on the
grid theta_j = pi (2j+1) / (4 LMAX), phi_i = 2 pi i / (2 LMAX).
"""
import ctypes as C

import numpy as np

from . import HIP_init, SphericalHarmonicsVolume, check, lib


def _vol(v):
    if not isinstance(v, SphericalHarmonicsVolume):
        raise TypeError("a SphericalHarmonicsVolume is required")
    return np.ascontiguousarray(v.coeff, dtype=np.complex128)


def getSurfaceVolume(volume, ctx=None):
    ctx = ctx or HIP_init()
    c = _vol(volume)
    out = np.zeros((volume.KMAX, 4 * volume.LMAX ** 2), np.complex128)
    check(lib().tp_sh_surface(ctx.h, volume.LMAX, volume.KMAX, c.ctypes.data, c.shape[1], out.ctypes.data),
          "tp_sh_surface")
    return out


def getSphericalHarmonicsVolume(surf, LMAX, KMAX, rmax, ctx=None):
    ctx = ctx or HIP_init()
    f = np.ascontiguousarray(surf, dtype=np.complex128).reshape(KMAX, 4 * LMAX * LMAX)
    out = np.zeros((KMAX, LMAX * LMAX), np.complex128)
    check(lib().tp_sh_coeff(ctx.h, LMAX, KMAX, f.ctypes.data, out.ctypes.data, LMAX * LMAX), "tp_sh_coeff")
    return SphericalHarmonicsVolume(out, LMAX, KMAX, rmax)


def _bessel(volume, fwd, ctx):
    ctx = ctx or HIP_init()
    c = _vol(volume)
    out = np.zeros((volume.KMAX, volume.LMAX ** 2), np.complex128)
    check(lib().tp_sph_bessel(ctx.h, volume.LMAX, volume.KMAX, int(fwd), c.ctypes.data, c.shape[1], out.ctypes.data),
          "tp_sph_bessel")
    # newvol.rmax = qmax(volume) (spherical_harmonics.jl:356)
    return SphericalHarmonicsVolume(out, volume.LMAX, volume.KMAX, np.pi / (volume.rmax / volume.KMAX))


def forward(volume, ctx=None):
    return _bessel(volume, True, ctx)


def backward(volume, ctx=None):
    return _bessel(volume, False, ctx)


def sRAAR(intensity, iterations, beta0=0.75, beta_max=0.75, tau=100.0, cutoff_factor=0.5, phases=None, seed=0,
          return_diff=False, ctx=None):
    """The density retrieved from an intensity volume.  phases: initial phases
    (KMAX, (2 LMAX)^2) in the surface sample order (the reference draws them
    with rand); None draws them from `seed`.  return_diff adds the
    per-iteration real-space change sum|density - tmp_u| / sum|density|."""
    ctx = ctx or HIP_init()
    c = _vol(intensity)
    K, LMAX = intensity.KMAX, intensity.LMAX
    ph = None
    if phases is not None:
        ph = np.ascontiguousarray(phases, dtype=np.float64).reshape(K, 4 * LMAX * LMAX)
    out = np.zeros((K, LMAX * LMAX), np.complex128)
    diff = np.zeros(max(iterations, 1))
    check(lib().tp_sraar(ctx.h, LMAX, K, c.ctypes.data, c.shape[1], int(iterations), float(beta0), float(beta_max),
                         float(tau), float(cutoff_factor), ph.ctypes.data if ph is not None else None, int(seed),
                         out.ctypes.data, diff.ctypes.data), "tp_sraar")
    vol = SphericalHarmonicsVolume(out, LMAX, K, np.pi / (intensity.rmax / K))
    return (vol, diff[:iterations]) if return_diff else vol
