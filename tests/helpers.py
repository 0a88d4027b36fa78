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


def quat_to_rot(q):
    """Rotation matrix of a unit quaternion (w, x, y, z)."""
    w, x, y, z = np.asarray(q, float) / np.linalg.norm(q)
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def sh_eval(oracle, c, l, LMAX, v):
    """sum_m c_{l,m} Y_lm(v) for unit vector v, Y_lm = sphPlm(l,m,theta) e^{i m phi}
    (sphPlm of the oracle: the reference's spherical_harmonics.jl:69-76)."""
    th, ph = np.arccos(np.clip(v[2], -1, 1)), np.arctan2(v[1], v[0])
    return sum(c[seanindex(l, m, LMAX) - 1] * oracle.sphPlm(l, m, th) * np.exp(1j * m * ph) for m in range(-l, l + 1))


def c2_invariants(c, K, L, LMAX):
    """sum_m conj(c_{k1,l,m}) c_{k2,l,m} for all k1 <= k2 <= K and even l <= L."""
    out = []
    for l in range(0, L + 1, 2):
        idx = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        blk = c[:K, idx]
        out.append(np.conj(blk) @ blk.T)
    return np.array(out)


def random_quats(rng, n):
    q = rng.normal(size=(n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)
