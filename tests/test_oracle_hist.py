"""Pins of the oracle's photon-pair / photon-triplet histograms
(or_hist_pictures: histogramCorrelationsInPicture_alltoall, datagen.jl:228-282)
against an independent numpy restatement (angles from atan2 differences
instead of acos and the cross product), hand cases (the orthogonal-photon
quirk of angle_between, utilities.jl:61-72; the alpha > pi flip; the k filter
and Julia's ties-to-even rounding) and combinatorial totals.  CPU only."""
import math

import numpy as np
import pytest


def _pictures(rng, npic, nmax, qcut, edge_free=True, N=8):
    """Random photons in the detector disk; angles kept away from bin edges
    and from exact orthogonality so that both restatements bin identically."""
    ph, off = [], [0]
    da = math.pi / N
    for _ in range(npic):
        n = int(rng.integers(0, nmax + 1))
        for _ in range(n):
            r = qcut * math.sqrt(rng.uniform(0.0, 1.0))
            if edge_free:
                t = (int(rng.integers(0, 4 * N)) + rng.uniform(0.15, 0.85)) * (da / 2)
            else:
                t = rng.uniform(0.0, 2 * math.pi)
            ph.append((r * math.cos(t), r * math.sin(t), 0.0))
        off.append(len(ph))
    return np.array(ph, dtype=np.float64).reshape(-1, 3), np.array(off, dtype=np.int64)


def _numpy_hist(ph, off, dq, N, K2, K3):
    """Independent restatement: angle_between(p, q) = (atan2 p - atan2 q) mod 2 pi."""
    da = math.pi / N
    n1 = np.zeros(K2, np.uint64)
    n2 = np.zeros((N, K2, K2), np.uint64)
    n3 = np.zeros((N, 2 * N, K3, K3, K3), np.uint64)
    for a in range(len(off) - 1):
        p = ph[off[a]:off[a + 1]]
        nr = np.sqrt((p * p).sum(axis=1))
        k = np.rint(nr / dq).astype(int)
        keep = (k >= 1) & (k <= K2)
        p, nr, k = p[keep], nr[keep], k[keep]
        if len(p) < 3:
            continue
        o = np.argsort(nr, kind="stable")
        p, k = p[o], k[o]
        th = np.arctan2(p[:, 1], p[:, 0])
        ang = lambda i, j: (th[i] - th[j]) % (2 * math.pi)
        for i in range(len(p)):
            n1[k[i] - 1] += 1
            for j in range(i):
                al = ang(i, j)
                als = 2 * math.pi - al if al > math.pi else al
                ai = int(math.floor(als / da)) % N
                n2[ai, k[j] - 1, k[i] - 1] += 1
                for m in range(j):
                    if k[i] <= K3 and k[j] <= K3 and k[m] <= K3:
                        be = ang(i, m)
                        if al > math.pi:
                            be = 2 * math.pi - be
                        bi = int(math.floor(be / da)) % (2 * N)
                        n3[ai, bi, k[m] - 1, k[j] - 1, k[i] - 1] += 1
    return n1, n2, n3


def test_hist_matches_numpy_restatement(oracle):
    rng = np.random.default_rng(3)
    N, K2, K3, qcut = 8, 12, 9, 1.0
    dq = qcut / K2
    ph, off = _pictures(rng, 12, 40, qcut, N=N)
    c1, c2, c3, n1, n2, n3 = oracle.hist_pictures(ph, off, dq, N, K2, K3)
    m1, m2, m3 = _numpy_hist(ph, off, dq, N, K2, K3)
    assert np.array_equal(n1, m1) and np.array_equal(n2, m2) and np.array_equal(n3, m3)
    assert n3.sum() > 1000
    # weights: every contribution of a bin has the same weight
    k = np.arange(1, K2 + 1)
    assert np.allclose(c1, n1 / k, rtol=1e-13, atol=0)
    for k1 in range(1, K3 + 1):
        for k2 in range(1, k1 + 1):
            w2 = 1.0 / ((2.0 if k1 != k2 else 1.0) * k1 * k2)
            assert np.allclose(c2[:, k2 - 1, k1 - 1], n2[:, k2 - 1, k1 - 1] * w2, rtol=1e-13, atol=0)
            for k3 in range(1, k2 + 1):
                w3 = 1.0 / (oracle.tripletFactor(k1, k2, k3) * k1 * k2 * k3)
                assert np.allclose(c3[:, :, k3 - 1, k2 - 1, k1 - 1], n3[:, :, k3 - 1, k2 - 1, k1 - 1] * w3,
                                   rtol=1e-13, atol=0)


def test_hist_totals(oracle):
    """Every photon, pair and (K3-limited) triplet of a picture is counted once."""
    rng = np.random.default_rng(5)
    N, K2, K3, qcut = 4, 10, 6, 1.0
    dq = qcut / K2
    ph, off = _pictures(rng, 1, 30, qcut, edge_free=False, N=N)
    _, _, _, n1, n2, n3 = oracle.hist_pictures(ph, off, dq, N, K2, K3)
    k = np.rint(np.sqrt((ph * ph).sum(axis=1)) / dq).astype(int)
    kept = k[(k >= 1) & (k <= K2)]
    l, l3 = len(kept), int(((kept >= 1) & (kept <= K3)).sum())
    assert n1.sum() == l and n2.sum() == l * (l - 1) // 2 and n3.sum() == l3 * (l3 - 1) * (l3 - 2) // 6


def test_hist_hand_cases(oracle):
    dq, N, K2, K3 = 1.0, 4, 4, 4
    # angle_between: orthogonal photons give 0 (not pi/2); cross_z > 0 flips to 2 pi - acos
    assert oracle.angle_between([2.0, 0, 0], [0, 3.0, 0]) == 0.0
    assert oracle.angle_between([1.0, 0, 0], [1.0, 1.0, 0]) == pytest.approx(2 * math.pi - math.pi / 4, rel=1e-15)
    assert oracle.angle_between([1.0, 1.0, 0], [1.0, 0, 0]) == pytest.approx(math.pi / 4, rel=1e-15)
    # k filter and ties to even: |p| = 2.5 -> k = 2; 0.4 -> 0 (dropped); 4.4 -> 4; 4.6 -> 5 (dropped)
    c = math.cos(0.3)
    s = math.sin(0.3)
    ph = np.array([[2.5, 0, 0], [0.4, 0, 0], [4.4 * c, 4.4 * s, 0], [4.6, 0.0, 0], [3.0 * c, -3.0 * s, 0]])
    c1, c2, c3, n1, n2, n3 = oracle.hist_pictures(ph, np.array([0, len(ph)]), dq, N, K2, K3)
    assert list(n1) == [0, 1, 1, 1]
    assert c1[1] == 0.5 and c1[2] == 1.0 / 3.0 and c1[3] == 0.25
    # sorted: (2.5 at 0) k=2, (3 at -0.3) k=3, (4.4 at 0.3) k=4.  Pair (k=4, k=2): cross_z < 0, so the
    # angle is acos = 0.3 -> ai = 0 (da = pi/4).  Triplet (4, 3, 2): alpha = angle(p_4.4, p_3) = 0.6
    # (< pi, no flip) -> ai = 0; beta = angle(p_4.4, p_2.5) = 0.3 -> bi = 0.
    assert n2[0, 1, 3] == 1 and n2.sum() == 3
    assert n3[0, 0, 1, 2, 3] == 1 and n3.sum() == 1
    assert c3[0, 0, 1, 2, 3] == pytest.approx(1.0 / (6.0 * 4 * 3 * 2), rel=1e-15)
