"""numpy restatement of the spherical-harmonics phase retrieval (sRAAR,
phases.jl:31-95) and of the transforms it uses.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the cpu_baseline leg of
bench.py -- never by the product package.

  surface            getSurfaceVolume per shell (spherical_harmonics.jl:139-153):
                     the series evaluated point by point with scipy's sph_harm_y
                     on the (2 LMAX)^2 grid (c2_oracle.surface)
  coefficients       getSphericalHarmonicsVolume per shell (:127-160): quadrature
                     with theta weights obtained by solving the Legendre moment
                     conditions sum_j w_j P_n(cos theta_j) = int P_n, n < 2 LMAX
                     (the Driscoll-Healy weights, found here without their
                     closed form)
  sph_bessel         sphBesselTransform (:296-377): frdr weights, Bessel J_{l+1/2}
                     from scipy.special.jv, (+-i)^l sqrt(1/k)
  sraar              phases.jl:6-95 with the initial phases given

Pinned in tests/test_oracle_phase.py: the weights integrate polynomials of
degree < 2 LMAX exactly and match the Driscoll-Healy closed form; analysis
inverts synthesis on band-limited shells; the Bessel transforms approach
their continuous limits (forward: i^l (8 pi / sqrt K) int r^2 j_l(q_k r) f dr,
backward: (-i)^l (2 / (pi sqrt K)) int q^2 j_l(q r) f dq, from the frdr
weights 8 j and J_{l+1/2}(x) = sqrt(2x/pi) j_l(x)) on smooth profiles, which
fixes the grids, the transposition, the powers of i and the sqrt(1/k)
factors; sRAAR's support and positivity constraints.
"""
import numpy as np
from scipy.special import eval_legendre, jv

from .c2_oracle import _ylm_table, surface
from .pyoracle import seanindex

__all__ = ["surface", "coefficients", "dh_weights", "sph_bessel", "sraar"]


def thetas(LMAX):
    return np.pi * (2 * np.arange(2 * LMAX) + 1) / (4 * LMAX)


def dh_weights(LMAX):
    """w_j with sum_j w_j f(cos theta_j) = int_{-1}^{1} f for polynomials of degree < 2 LMAX."""
    x = np.cos(thetas(LMAX))
    n = 2 * LMAX
    A = np.stack([eval_legendre(k, x) for k in range(n)])
    b = np.zeros(n)
    b[0] = 2.0
    return np.linalg.solve(A, b)


def coefficients(f, LMAX):
    """(LMAX^2,) seanindex-layout coefficients of one shell's (2 LMAX)^2 surface."""
    S = 2 * LMAX
    w = dh_weights(LMAX)
    Y = _ylm_table(LMAX)
    wf = np.asarray(f).reshape(S, S) * w[:, None] * (2 * np.pi / S)
    return np.tensordot(np.conj(Y), wf, axes=([1, 2], [0, 1]))


def _frdr(x):
    n = len(x)
    res = np.empty(n)
    res[0] = np.sqrt(0.5) * (x[0] + x[1]) ** 2
    for j in range(1, n - 1):
        res[j] = (x[j + 1] - x[j - 1]) * (x[j - 1] + 2 * x[j] + x[j + 1])
    res[n - 1] = 4 * x[n - 1] ** 2 - (x[n - 2] + x[n - 1]) ** 2
    return np.pi / 4 * res ** 1.5


def sph_bessel(coeff, LMAX, forward):
    """sphBesselTransform of (K, >= LMAX^2) coefficients (shell k = row k-1)."""
    c = np.asarray(coeff, dtype=np.complex128)
    K = c.shape[0]
    r = np.arange(1.0, K + 1)
    q = np.pi / K * r
    out = np.zeros((K, LMAX * LMAX), np.complex128)
    for l in range(LMAX):
        I = jv(l + 0.5, q[:, None] * r[None, :])          # I[k][r]
        fac = (1j) ** l if forward else (-1j) ** l
        for m in range(-l, l + 1):
            s = seanindex(l, m, LMAX) - 1
            rho = c[:, s]
            if forward:
                v = I @ (_frdr(r) * rho)
            else:
                v = I.T @ (_frdr(q) * rho) / (2 * np.pi) ** 2
            out[:, s] = fac * np.sqrt(1.0 / r) * v
    return out


def _forward_surf(surf, LMAX):
    c = np.array([coefficients(s, LMAX) for s in surf])
    c = sph_bessel(c, LMAX, True)
    return np.array([surface(x, LMAX).ravel() for x in c])


def _backward_surf(surf, LMAX):
    c = np.array([coefficients(s, LMAX) for s in surf])
    c = sph_bessel(c, LMAX, False)
    return np.array([surface(x, LMAX).ravel() for x in c])


def _A(x, Koff):
    out = np.maximum(x.real, 0.0).astype(np.complex128)
    out[Koff:] = 0.0
    return out


def _B(x, amp, LMAX):
    F = _forward_surf(x, LMAX)
    F = amp * np.exp(1j * np.angle(F))
    return _backward_surf(F, LMAX)


def sraar(intensity, LMAX, iterations, beta0, beta_max, tau, cutoff_factor, phases):
    """(coefficients (K, LMAX^2), diff list) of sRAAR with the given initial
    phases (K, (2 LMAX)^2)."""
    I = np.asarray(intensity, dtype=np.complex128)
    K = I.shape[0]
    Koff = int(np.ceil(cutoff_factor * K))
    amp = np.array([np.sqrt(np.maximum(surface(I[k], LMAX).real.ravel(), 0.0)) for k in range(K)])
    F = amp * np.exp(1j * np.asarray(phases))
    density = _backward_surf(F, LMAX)
    tmp1 = 2.0 * _B(density, amp, LMAX) - density
    tmp2 = tmp1
    diffs = []
    for n in range(1, iterations + 1):
        e3 = np.exp((-n / tau) ** 3.0)
        beta = e3 * beta0 + (1.0 - e3) * beta_max
        tmp3 = _A(tmp1, Koff)
        tmp_u = 0.5 * (beta * (2.0 * tmp3 - tmp1) + (1.0 - beta) * tmp1 + density)
        tmp2 = _B(tmp_u, amp, LMAX)
        diffs.append(np.abs(density - tmp_u).sum() / np.abs(density).sum())
        density = tmp_u
        tmp1 = 2.0 * tmp2 - tmp_u
    final = _A(tmp2, Koff)   # with no iteration tmp2 is tmp1 (phases.jl:59)
    return np.array([coefficients(s, LMAX) for s in final]), diffs
