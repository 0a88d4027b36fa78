"""numpy restatement of the two-photon correlation and of retrieveSolution.

TEST INFRASTRUCTURE ONLY: imported by tests/ and by the cpu_baseline leg of
bench.py -- never by the product package.

  two_photons        twoPhotons, correlations.jl:229-243 (lambda = 0, so
                     alpha_star(alpha) = alpha, datagen.jl:49-53; alpharange,
                     correlations.jl:247-250)
  retrieve_solution  retrieveSolution, correlations.jl:262-319, with
                     real_to_comp (spherical_harmonics.jl:472-481) and
                     negativityCheck (spherical_harmonics.jl:625-628)

Independent of the library's arithmetic: Legendre polynomials come from
scipy.special.eval_legendre, the fits from numpy.linalg.lstsq (LAPACK), the
eigenpairs from numpy.linalg.eigh and the surfaces from scipy's sph_harm_y.
The eigenvector sign, which LAPACK leaves open (the reference inherits
whatever its LAPACK returns), is fixed the way the library fixes it: the
entry of largest magnitude is positive.  Pinned in tests/test_oracle_c2.py by
closed forms (an l = 0 structure, a single-l structure), by the invariance of
c2 under rotations of the l-blocks, by the round trip c2 -> structure -> c2
and by the rank of G.
"""
import numpy as np
from scipy.special import eval_legendre, sph_harm_y

from .pyoracle import seanindex, umat


def alpharange(N):
    """correlations.jl:247-250: N points from da/2 to pi - da/2, da = pi/N."""
    da = np.pi / N
    return np.linspace(da / 2, np.pi - da / 2, N)


def alpha_star(alpha, k1, k2, dq, lam):
    """datagen.jl:49-53: the angle between the Ewald-sphere points of shells k1,
    k2 (1-based) whose detector azimuths differ by alpha."""
    t1 = np.arccos(k1 * dq * lam / (4 * np.pi))
    t2 = np.arccos(k2 * dq * lam / (4 * np.pi))
    return np.arccos(np.sin(t1) * np.sin(t2) * np.cos(alpha) + np.cos(t1) * np.cos(t2))


def two_photons(coeff, L, LMAX, N, K2, minimal=False, normalize=False, lam=0.0, dq=0.0):
    """c2 (N, K2, K2) indexed [a, k2, k1] (zero-based) from coefficients
    coeff (>= K2, LMAX^2) complex in the seanindex layout; lam != 0 takes the
    Legendre series at alpha_star with the volume's shell spacing dq."""
    c = np.asarray(coeff, dtype=np.complex128)
    al = alpharange(N)
    out = np.zeros((N, K2, K2))
    for k1 in range(K2):
        for k2 in range(k1 + 1 if minimal else K2):
            sl = np.zeros(N, np.complex128)
            for l in range(0, L + 1, 2):
                ix = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
                fac = np.vdot(c[k1, ix], c[k2, ix])       # sum conj(a) b
                x = np.cos(al) if lam == 0.0 else np.cos(alpha_star(al, k1 + 1, k2 + 1, dq, lam))
                sl = sl + fac * eval_legendre(l, x)
            out[:, k2, k1] = sl.real
    return out / np.abs(out).sum() if normalize else out / (4 * np.pi)


def _fix_sign(V):
    imax = np.argmax(np.abs(V), axis=0)
    s = np.sign(V[imax, np.arange(V.shape[1])])
    s[s == 0] = 1.0
    return V * s


_YTAB = {}


def _ylm_table(LMAX):
    """Ylm at every point of the (2 LMAX)^2 grid, rows in seanindex order (cached)."""
    if LMAX not in _YTAB:
        S = 2 * LMAX
        th = np.pi * (2 * np.arange(S) + 1) / (4 * LMAX)
        ph = 2 * np.pi * np.arange(S) / S
        T, P = np.meshgrid(th, ph, indexing="ij")
        Y = np.zeros((LMAX * LMAX, S, S), np.complex128)
        for l in range(LMAX):
            for m in range(-l, l + 1):
                Y[seanindex(l, m, LMAX) - 1] = sph_harm_y(l, m, T, P)
        _YTAB[LMAX] = Y
    return _YTAB[LMAX]


def surface(c, LMAX):
    """Values of one shell's series sum c_lm Ylm on the (2 LMAX)^2 grid
    theta_j = pi (2j+1)/(4 LMAX), phi_i = 2 pi i/(2 LMAX) (Ylm from scipy's
    sph_harm_y at each grid point)."""
    Y = _ylm_table(LMAX)
    return np.tensordot(np.asarray(c[:LMAX * LMAX], dtype=np.complex128), Y, axes=(0, 0))


def negativity(coeff, LMAX, lo, hi):
    """negativityCheck over shells lo..hi (1-based)."""
    tot = 0.0
    for k in range(lo, hi + 1):
        f = surface(coeff[k - 1], LMAX)
        with np.errstate(invalid="ignore", divide="ignore"):
            r = np.abs(np.minimum(f.real, 0.0)).sum() / np.abs(f).sum()
        tot += r if np.isnan(r) else max(r, 0.0)
    return tot


def retrieve_solution(c2, L, LMAX, lo, hi, qmax=1.0, lam=0.0):
    """(coeff (hi, LMAX^2), info): coefficients fitted to c2 ([a, k2, k1]) over
    shells lo..hi (1-based); info holds G and the eigenvalues per l, the
    negativity before the flip and whether every sign was changed.  lam != 0:
    Legendre columns at alpha_star with dq = qmax / hi, no symmetrisation."""
    N = c2.shape[0]
    K2 = hi - lo + 1
    assert 2 * L < K2
    al = alpharange(N)
    out = np.zeros((hi, LMAX * LMAX), np.complex128)
    info = {"G": {}, "w": {}}
    for l in range(0, L + 1, 2):
        G = np.zeros((K2, K2))
        for k1 in range(lo, hi + 1):
            for k2 in range(lo, k1 + 1):
                s = c2[:, k2 - 1, k1 - 1]
                if lam == 0.0:
                    s = 0.5 * (s + s[::-1])
                upperL = min(L, 2 * min(k1, k2) - 2)
                x = np.cos(al) if lam == 0.0 else np.cos(alpha_star(al, k1, k2, qmax / hi, lam))
                A = np.stack([eval_legendre(lp, x) / (4 * np.pi) for lp in range(0, upperL + 1, 2)], axis=1)
                fac = np.linalg.lstsq(A, s, rcond=None)[0]
                if l <= upperL:
                    G[k2 - lo, k1 - lo] = G[k1 - lo, k2 - lo] = fac[l // 2]
        w, V = np.linalg.eigh(G)
        n = 2 * l + 1
        w, V = w[K2 - n:], _fix_sign(V[:, K2 - n:])
        x = np.sqrt(np.maximum(0.0, w))[:, None] * V.T       # (2l+1, K2), row j <-> m = j - l
        cl = umat(l).conj().T @ x                             # real_to_comp
        ix = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        for k in range(lo, hi + 1):
            out[k - 1, ix] = cl[:, k - lo]
        info["G"][l], info["w"][l] = G, w
    neg = negativity(out, LMAX, lo, hi)
    info["negativity"], info["flipped"] = neg, bool(neg > 0.33)
    if neg > 0.33:
        out = -out
    return out, info
