"""numpy restatement of rotateStructure and similarity.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the cpu_baseline leg of
bench.py -- never by the product package.

  rotate_structure  rotateStructure, spherical_harmonics.jl:598-612, with
                    rotationMatrix / Dlms / d_work (:572-595) from the C oracle
                    (or_rotation_matrix, the factorial sum of d_work)
  similarity        similarity, spherical_harmonics.jl:426-435: the mean over
                    shells of cor_ISC (utilities.jl:45-47, Pearson's correlation
                    of the real parts) of the shells' surfaces, each evaluated
                    point by point on the (2 LMAX)^2 grid with scipy's
                    sph_harm_y (c2_oracle.surface)

The device scorer sums over phi in closed form (Parseval on the uniform phi
grid) and rotates by quaternion Wigner-D recurrences; this file does neither.
Pinned in tests/test_oracle_fit.py: the rotation is f(R^-1 r) with
R = Rz(phi) Ry(-theta) Rz(gamma) at random points; similarity is 1 for a
structure with itself or with a positive affine image of it, -1 with its
negative, and below 1 for a rotated copy.
"""
import numpy as np

from .c2_oracle import surface
from .pyoracle import rotation_matrix, seanindex


def rotate_structure(coeff, theta, phi, gamma, K, lrange, LMAX):
    c = np.array(coeff, dtype=np.complex128, copy=True)
    for l in lrange:
        D = rotation_matrix(theta, phi, gamma, l)
        ix = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        c[:K, ix] = (D @ c[:K, ix].T).T
    return c


def similarity(reference, volume, LMAX, klo, khi):
    cs = []
    for k in range(klo, khi + 1):
        a = surface(reference[k - 1], LMAX).real.ravel()
        b = surface(volume[k - 1], LMAX).real.ravel()
        cs.append(np.corrcoef(a, b)[0, 1])
    return float(np.sum(cs) / len(cs))
