"""Shared helpers for the tests (random real-intensity coefficients, k-triple walks)."""
import numpy as np


def seanindex(l, m, LMAX):
    B = LMAX
    if m >= 0:
        return 1 + m * B - m * (m - 1) // 2 + (l - m)
    a = B - 1 + m
    return 1 + B * (B + 1) // 2 + a * (a + 1) // 2 + (l + m)


def random_real_coeff(K, L, LMAX, seed, even=True, c00=3.0, decay=0.0):
    """Coefficients of a real function: c_{l,-m} = (-1)^m conj(c_{l,m})."""
    rng = np.random.default_rng(seed)
    c = np.zeros((K, LMAX * LMAX), complex)
    for k in range(K):
        for l in range(0, L + 1, 2 if even else 1):
            amp = np.exp(-decay * l)
            for m in range(0, l + 1):
                v = amp * (rng.normal() + (1j * rng.normal() if m > 0 else 0))
                if l == 0:
                    v = v + c00
                c[k, seanindex(l, m, LMAX) - 1] = v
                if m > 0:
                    c[k, seanindex(l, -m, LMAX) - 1] = (-1) ** m * np.conj(v)
    return c


def ktriples(K, lo=1):
    return [(k1, k2, k3) for k1 in range(lo, K + 1) for k2 in range(lo, k1 + 1) for k3 in range(lo, k2 + 1)]


def c3ref_from_columns(cols, N, K):
    """Place R x klength columns into an (N, 2N, K, K, K) array indexed [a, b, k3, k2, k1]."""
    h = np.zeros((N, 2 * N, K, K, K))
    for i, (k1, k2, k3) in enumerate(ktriples(K)):
        h[:, :, k3 - 1, k2 - 1, k1 - 1] = cols[:, i].reshape(2 * N, N).T
    return h


def rel_err(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(np.asarray(b)).max(), 1e-300))
