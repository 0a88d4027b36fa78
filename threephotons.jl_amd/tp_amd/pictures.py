"""Synthetic sparse pictures on the device (tp_pictures; pointsPerOrientation,
datagen.jl:88-144) and the histogram generation over them without leaving
the device (tp_hist_add_generated; generateHistogram's loop, datagen.jl:391-415).

The intensity is a SphericalHarmonicsVolume; as in the reference's data
generation it is sampled through its surfaces (getSurfaceVolume) with
getVolumeInterpolated's trilinear interpolation.  Defaults follow
generateHistogram: envelope width qcut/3, 1000 incident photons, lambda 1.
"""
import ctypes as C

import numpy as np

from . import HIP_init, SphericalHarmonicsVolume, check, lib
from ._lib import PictureParams


def _params(qcut, envelope_sigma, photons_per_image, lam, beamstop_width, capacity, seed, particles=1,
            noise_photons=0, noise_sigma=1.0, incident_photon_variance=0, beamstop_only=False):
    return PictureParams(float(qcut), float(qcut / 3.0 if envelope_sigma is None else envelope_sigma), float(lam),
                         float(beamstop_width), int(photons_per_image), int(capacity), int(seed), int(particles),
                         int(noise_photons), float(noise_sigma), int(incident_photon_variance), int(bool(beamstop_only)))


def _coeff(volume):
    if not isinstance(volume, SphericalHarmonicsVolume):
        raise TypeError("a SphericalHarmonicsVolume is required")
    return np.ascontiguousarray(volume.coeff, dtype=np.complex128)


def pointsPerOrientation(volume, qcut, npictures, envelope_sigma=None, photons_per_image=1000, lam=1.0,
                         beamstop_width=0.0, capacity=512, seed=0, first_picture=0, particles=1, noise_photons=0,
                         noise_sigma=1.0, incident_photon_variance=0, beamstop_only=False, ctx=None):
    """(pictures, rotations): pictures[i] = (n_i, 3) accepted photons of picture
    first_picture + i (detector positions, z = 0), rotations (npictures, 3, 3)
    of the first particle.  particles / noise_photons / noise_sigma /
    incident_photon_variance / beamstop_only as generateHistogram's
    number_particles, GaussianNoise(gamma > 0, sigma, photons),
    incident_photon_variance and beamstop_only."""
    ctx = ctx or HIP_init()
    c = _coeff(volume)
    p = _params(qcut, envelope_sigma, photons_per_image, lam, beamstop_width, capacity, seed, particles,
                noise_photons, noise_sigma, incident_photon_variance, beamstop_only)
    ph = np.zeros((npictures, capacity, 3))
    cnt = np.zeros(npictures, np.int32)
    rot = np.zeros((npictures, 3, 3))
    check(lib().tp_pictures(ctx.h, volume.LMAX, volume.KMAX, volume.rmax, c.ctypes.data, c.shape[1], C.byref(p),
                            npictures, first_picture, ph.ctypes.data, cnt.ctypes.data, rot.ctypes.data), "tp_pictures")
    if np.any(cnt > capacity):
        raise RuntimeError(f"a picture accepted {cnt.max()} photons, above capacity {capacity}")
    return [ph[i, :cnt[i]].copy() for i in range(npictures)], rot


def generate_histograms(hist, volume, qcut, npictures, envelope_sigma=None, photons_per_image=1000, lam=1.0,
                        beamstop_width=0.0, capacity=512, seed=0, first_picture=0, particles=1, noise_photons=0,
                        noise_sigma=1.0, incident_photon_variance=0, beamstop_only=False):
    """Add npictures synthetic pictures of `volume` to the PhotonHistogram
    `hist`.  Raises RuntimeError when a picture accepted more photons than
    `capacity` (as pointsPerOrientation does); such pictures are left out of
    the histogram, the others are counted."""
    c = _coeff(volume)
    p = _params(qcut, envelope_sigma, photons_per_image, lam, beamstop_width, capacity, seed, particles,
                noise_photons, noise_sigma, incident_photon_variance, beamstop_only)
    over = C.c_int64()
    check(lib().tp_hist_add_generated(hist.h, volume.LMAX, volume.KMAX, volume.rmax, c.ctypes.data, c.shape[1],
                                      C.byref(p), npictures, first_picture, C.byref(over)), "tp_hist_add_generated")
    if over.value > 0:
        raise RuntimeError(f"{over.value} of {npictures} pictures accepted more photons than capacity {capacity} "
                           "and were left out of the histogram")
    return 0
