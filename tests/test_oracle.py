"""Pins of the CPU oracle (oracle/tp_oracle.c) against golden vectors, closed
forms, invariants and brute force.  CPU only."""
import json
import math
import os

import numpy as np
import pytest

from helpers import c3ref_from_columns, ktriples, random_real_coeff, rel_err, seanindex

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---------------------------------------------------------------- special functions
def test_wigner3j_golden(oracle):
    rows = json.load(open(os.path.join(GOLD, "wigner3j.json")))
    worst = 0.0
    for l1, m1, l2, m2, l3, m3, v in rows:
        got = oracle.wigner3j(l1, m1, l2, m2, l3, m3)
        worst = max(worst, abs(got - v) / max(abs(v), 1e-3))
    assert worst < 1e-14, worst


def test_wigner3j_closed_forms(oracle):
    for l in range(0, 19):
        for m in range(-l, l + 1):
            assert oracle.wigner3j(l, m, l, -m, 0, 0) == pytest.approx((-1) ** (l - m) / math.sqrt(2 * l + 1), rel=1e-14)
    # selection rules
    assert oracle.wigner3j(2, 1, 2, 1, 2, 0) == 0.0          # m1+m2+m3 != 0
    assert oracle.wigner3j(2, 0, 2, 0, 6, 0) == 0.0          # triangle
    assert oracle.wigner3j(1, 0, 1, 0, 1, 0) == 0.0          # odd l-sum with m = 0


def test_wigner3j_orthogonality(oracle):
    # sum_{m1,m2} (2l3+1) 3j(l1 l2 l3; m1 m2 m3) 3j(l1 l2 l3'; m1 m2 m3') = delta
    for (l1, l2) in [(4, 6), (10, 8), (18, 18)]:
        top = min(l1 + l2, 22)   # the exact-sum 3j covers l <= 22
        for l3 in (abs(l1 - l2), (abs(l1 - l2) + top) // 2, top):
            for l3p in (l3, min(l3 + 2, top)):
                for m3 in (0, min(l3, l3p)):
                    s = 0.0
                    for m1 in range(-l1, l1 + 1):
                        m2 = -m1 - m3
                        if abs(m2) > l2:
                            continue
                        s += oracle.wigner3j(l1, m1, l2, m2, l3, m3) * oracle.wigner3j(l1, m1, l2, m2, l3p, m3)
                    want = 1.0 / (2 * l3 + 1) if l3 == l3p else 0.0
                    assert abs(s - want) < 1e-14


def test_sphplm_golden(oracle):
    rows = json.load(open(os.path.join(GOLD, "sphplm.json")))
    worst = max(abs(oracle.sphPlm(l, m, th) - v) / max(abs(v), 1e-2) for l, m, th, v in rows)
    assert worst < 1e-13, worst


def test_sphplm_closed_forms(oracle):
    th = math.acos(0.0)
    for l in range(0, 19):
        for m in range(0, l + 1):
            got = oracle.sphPlm(l, m, th)
            if (l + m) % 2:
                assert abs(got) < 1e-15
            else:
                # P_l^m(0) = (-1)^((l+m)/2) (l+m-1)!! / (l-m)!!  (with Condon-Shortley phase)
                df = lambda n: 1 if n <= 0 else n * df(n - 2)
                p = (-1) ** ((l + m) // 2) * df(l + m - 1) / df(l - m)
                want = math.sqrt((2 * l + 1) / (4 * math.pi) * math.factorial(l - m) / math.factorial(l + m)) * p
                assert got == pytest.approx(want, rel=1e-13, abs=1e-15)
            # negative m: (-1)^m (spherical_harmonics.jl:73-75)
            assert oracle.sphPlm(l, -m, 1.1) == pytest.approx((-1) ** m * oracle.sphPlm(l, m, 1.1), rel=1e-15)
    # addition theorem: sum_m |Y_lm|^2 = (2l+1)/(4 pi)
    for l in (3, 8, 18):
        s = sum(oracle.sphPlm(l, m, 0.7) ** 2 for m in range(-l, l + 1))
        assert s == pytest.approx((2 * l + 1) / (4 * math.pi), rel=1e-13)


def test_seanindex_bijection(oracle):
    for LMAX in (1, 5, 25):
        seen = sorted(oracle.seanindex(l, m, LMAX) for l in range(LMAX) for m in range(-l, l + 1))
        assert seen == list(range(1, LMAX * LMAX + 1))
        for l in range(LMAX):
            for m in range(-l, l + 1):
                assert oracle.seanindex(l, m, LMAX) == seanindex(l, m, LMAX)


def test_triplet_factor(oracle):
    assert oracle.tripletFactor(3, 3, 3) == 1.0
    assert oracle.tripletFactor(3, 3, 1) == 3.0
    assert oracle.tripletFactor(3, 1, 1) == 3.0
    assert oracle.tripletFactor(3, 2, 1) == 6.0


def test_philox_kat(oracle):
    kat = json.load(open(os.path.join(GOLD, "philox_kat.json")))
    for e in kat:
        assert list(map(int, oracle.philox(e["ctr"], e["key"]))) == e["out"]
    # Random123 kat_vectors, philox4x32_10
    assert list(oracle.philox([0, 0, 0, 0], [0, 0])) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert list(oracle.philox([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0])) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


# ---------------------------------------------------------------- basis
@pytest.mark.parametrize("L,basislen,n_lt,n_m", [(8, 4583, 65, 217), (18, 163153, 505, 1027)])
def test_basis_counts(oracle, L, basislen, n_lt, n_m):
    b = oracle.Basis(L, 25, 4, 2)
    assert (b.basislen, b.n_lt, b.n_m) == (basislen, n_lt, n_m)


def test_basis_layout(oracle):
    b = oracle.Basis(4, 6, 3, 3)
    w, ind, lt = b.tables()
    for j in range(b.basislen):
        l1, m1, l2, m2, l3, mu3 = ind[j, 3:]
        assert m1 + m2 + mu3 == 0
        assert ind[j, 0] == seanindex(l1, m1, 6) and ind[j, 1] == seanindex(l2, m2, 6) and ind[j, 2] == seanindex(l3, mu3, 6)
        assert w[j] == oracle.wigner3j(l1, m1, l2, m2, l3, mu3)
    assert lt[:, 3].sum() == b.basislen and lt[0, 4] == 1
    pc = b.pacombos()
    assert pc.shape == (b.klength * b.n_lt, 9)
    assert tuple(pc[0]) == (1, 1, 1, 1, 0, 0, 0, 1, 1)


# ---------------------------------------------------------------- c3
def _tiny(seed=0, L=4, LMAX=6, N=4, K=3):
    return random_real_coeff(K, L, LMAX, seed), L, LMAX, N, K


def test_c3_three_formulations_agree(oracle):
    c, L, LMAX, N, K = _tiny()
    b = oracle.Basis(L, LMAX, N, K)
    d = b.c3_dense(c)
    f = b.c3_factored(c)
    s = np.stack([oracle.c3_slice(c, N, L, LMAX, *kt) for kt in ktriples(K)], axis=1)
    assert rel_err(f, d) < 1e-13
    assert rel_err(s, d) < 1e-13


def test_c3_curved_ewald_dense_equals_slice(oracle):
    """lambda != 0: the dense B.PA form (Legendre factors at theta_k =
    acos(k lambda dq / 4 pi), odd m included) equals the slow flab form."""
    c, L, LMAX, N, K = _tiny(seed=4)
    lam, dq = 2.5, 1.0
    b = oracle.Basis(L, LMAX, N, K, lam=lam, dq=dq)
    d = b.c3_dense(c)
    s = np.stack([oracle.c3_slice(c, N, L, LMAX, *kt, dq=dq, lam=lam) for kt in ktriples(K)], axis=1)
    assert rel_err(s, d) < 1e-13
    flat = oracle.Basis(L, LMAX, N, K).c3_dense(c)
    assert rel_err(d, flat) > 1e-3   # the curvature changes c3


def test_c3_small_config_dense_vs_factored(oracle):
    c = random_real_coeff(8, 8, 25, 3, decay=0.2)
    b = oracle.Basis(8, 25, 16, 8)
    d = b.c3_dense(c, 0, 20)
    f = b.c3_factored(c)[:, :20]
    assert rel_err(f, d) < 1e-12


def _Y(oracle, l, m, th, ph):
    return oracle.sphPlm(l, m, th) * np.exp(1j * m * ph)


@pytest.mark.parametrize("lam,dq", [(0.0, 0.0), (2.5, 1.0)])
def test_c3_equals_orientational_average(oracle, lam, dq):
    """c3(a,b) = (1/8pi^2) int_SO(3) I(R n1) I(R n2) I(R n3) dR, n_i at polar
    angles theta_ki = acos(k_i lambda dq / 4 pi) (the Ewald sphere of
    detector_to_Ewald_sphere, datagen.jl:33-39; pi/2 for lambda = 0) and
    azimuths 0, alpha, beta; exact quadrature over Euler angles for a
    band-limited intensity."""
    c, L, LMAX, N, K = _tiny(seed=5)
    b = oracle.Basis(L, LMAX, N, K, lam=lam, dq=dq)
    d = b.c3_dense(c)

    def I(k, n):
        x, y, z = n
        th, ph = math.acos(max(-1, min(1, z))), math.atan2(y, x)
        v = 0
        for l in range(0, L + 1, 2):
            for m in range(-l, l + 1):
                v += c[k - 1, seanindex(l, m, LMAX) - 1] * _Y(oracle, l, m, th, ph)
        return v.real

    def Rz(a):
        return np.array([[math.cos(a), -math.sin(a), 0], [math.sin(a), math.cos(a), 0], [0, 0, 1]])

    def Ry(a):
        return np.array([[math.cos(a), 0, math.sin(a)], [0, 1, 0], [-math.sin(a), 0, math.cos(a)]])

    gx, gw = np.polynomial.legendre.leggauss(8)
    na = 14
    kts = ktriples(K)
    for kt, ia, jb in [((3, 2, 2), 1, 5), ((2, 2, 1), 0, 7), ((3, 3, 3), 3, 2)]:
        a, bb = (ia + 0.5) * math.pi / N, (jb + 0.5) * math.pi / N
        th = [math.acos(k * lam * dq / (4 * math.pi)) for k in kt]
        n1, n2, n3 = (np.array([math.sin(t) * math.cos(p), math.sin(t) * math.sin(p), math.cos(t)])
                      for t, p in zip(th, (0.0, a, bb)))
        acc = 0.0
        for xb, wb in zip(gx, gw):
            for ai in range(na):
                for gi in range(na):
                    Rm = Rz(2 * math.pi * ai / na) @ Ry(math.acos(xb)) @ Rz(2 * math.pi * gi / na)
                    acc += wb * I(kt[0], Rm @ n1) * I(kt[1], Rm @ n2) * I(kt[2], Rm @ n3)
        acc *= (2 * math.pi / na) ** 2 / (8 * math.pi ** 2)
        assert d[ia + jb * N, kts.index(kt)] == pytest.approx(acc, rel=1e-11, abs=1e-12)


def test_c3_so3_invariance(oracle):
    c, L, LMAX, N, K = _tiny(seed=9, L=6, LMAX=8, N=5, K=3)
    b = oracle.Basis(L, LMAX, N, K)
    base = b.c3_factored(c)
    rot = c.copy()
    for l in range(0, L + 1, 2):
        D = oracle.rotation_matrix(0.7, 1.9, -0.4, l)
        idx = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        rot[:, idx] = c[:, idx] @ D.T
    assert rel_err(b.c3_factored(rot), base) < 1e-12


def test_c3_single_shell_closed_form(oracle):
    L, LMAX, N, K = 6, 8, 4, 2
    c = np.zeros((K, LMAX * LMAX), complex)
    c[0, 0], c[1, 0] = 2.0, -1.5
    b = oracle.Basis(L, LMAX, N, K)
    d = b.c3_dense(c)
    for i, (k1, k2, k3) in enumerate(ktriples(K)):
        want = c[k1 - 1, 0].real * c[k2 - 1, 0].real * c[k3 - 1, 0].real / (4 * math.pi) ** 1.5
        assert np.allclose(d[:, i], want, rtol=1e-14, atol=0)


def test_c3_lambda0_parity_zeros(oracle):
    # at theta = pi/2 only even-m entries of F carry weight (P_l^m(0) = 0 for l+m odd)
    b = oracle.Basis(6, 8, 4, 2)
    w, ind, lt = b.tables()
    for j in range(0, b.basislen, 7):
        l1, m1, l2, m2, l3, mu3 = ind[j, 3:]
        y = oracle.sphPlm(l1, m1, math.pi / 2) * oracle.sphPlm(l2, m2, math.pi / 2) * oracle.sphPlm(l3, mu3, math.pi / 2)
        if m1 % 2:
            assert abs(y) < 1e-15


def test_c3_fp32_tables_gap(oracle):
    """The reference stores B and P as Float32 (correlations.jl:100-101); the
    FP64-basis oracle differs from that by ~1e-7, the size of FP32 rounding."""
    c = random_real_coeff(3, 6, 8, 1)
    b64 = oracle.Basis(6, 8, 4, 3)
    b32 = oracle.Basis(6, 8, 4, 3, fp32_tables=True)
    e = rel_err(b32.c3_dense(c), b64.c3_dense(c))
    assert 1e-9 < e < 1e-5


def _orbits(N):
    """Bins r = i + N j of the orbits {(i, j), (i, j+N), (N-1-i, 2N-1-j),
    (N-1-i, N-1-j)} of the representatives i < N/2, j < N (even N)."""
    i, j = np.meshgrid(np.arange(N // 2), np.arange(N), indexing="ij")
    return [a.ravel() + N * b.ravel() for a, b in
            ((i, j), (i, j + N), (N - 1 - i, 2 * N - 1 - j), (N - 1 - i, N - 1 - j))]


def test_c3_lambda0_folded_grid(oracle):
    """lambda = 0: every class has even m2 and mu3, so on the midpoint grids
    c3 takes one value on each orbit of four bins, the orbits partition the
    2N^2 bins, and the energy from the N^2/2 representatives with the folded
    histogram (S: 4 p per bin; Q: the orbit's summed q) is the full-grid
    energy (the folding of the library's k_c3sep FOLD)."""
    L, LMAX, N, K = 8, 10, 8, 3
    c = random_real_coeff(K, L, LMAX, 21, c00=5.0)
    b = oracle.Basis(L, LMAX, N, K)
    c3 = b.c3_factored(c)
    orb = _orbits(N)
    allbins = np.sort(np.concatenate(orb))
    assert np.array_equal(allbins, np.arange(2 * N * N))
    scale = np.abs(c3).max()
    for o in orb[1:]:
        assert np.abs(c3[o] - c3[orb[0]]).max() <= 1e-13 * scale
    # the energy from the representatives and the folded histogram
    ref_cols = np.random.default_rng(5).uniform(0.1, 1.0, size=c3.shape)
    ref = c3ref_from_columns(ref_cols, N, K)
    E_full = _energy(oracle, c3, N, ref, K)
    E_fold = 0.0
    for i, (k1, k2, k3) in enumerate(ktriples(K)):
        q = ref_cols[:, i] / ref_cols[:, i].sum()
        p = np.maximum(c3[orb[0], i], 1e-30)
        qf = sum(q[o] for o in orb)
        S, Q = 4.0 * p.sum(), (qf * np.log(p)).sum()
        E_fold += oracle.tripletFactor(k1, k2, k3) * (q.sum() * math.log(S) - Q)
    assert E_fold == pytest.approx(E_full, rel=1e-13)


def test_c3_curved_ewald_no_folding(oracle):
    """lambda dq != 0 brings odd m2 and mu3: c3 differs across an orbit, so
    the library keeps the full grid there."""
    L, LMAX, N, K = 6, 8, 4, 2
    c = random_real_coeff(K, L, LMAX, 22, c00=5.0)
    c3 = oracle.Basis(L, LMAX, N, K, lam=2.5, dq=1.0).c3_dense(c)
    orb = _orbits(N)
    assert np.abs(c3[orb[1]] - c3[orb[0]]).max() > 1e-6 * np.abs(c3).max()


# ---------------------------------------------------------------- energy
def _energy(oracle, c3raw, N, c3ref, K, lo=1):
    return oracle.energy(c3raw, N, c3ref, c3ref.shape[2], lo, K)


def test_energy_hand_example(oracle):
    c3raw = np.array([[2.0], [6.0]])                 # N=1: R=2 bins, K=1: one k-triple
    ref = np.zeros((1, 2, 1, 1, 1))
    ref[0, :, 0, 0, 0] = [1.0, 3.0]
    want = -(0.25 * math.log(2 / 8) + 0.75 * math.log(6 / 8))
    assert _energy(oracle, c3raw, 1, ref, 1) == pytest.approx(want, rel=1e-15)


def test_energy_clamp(oracle):
    c3raw = np.array([[-5.0], [4.0]])
    ref = np.zeros((1, 2, 1, 1, 1))
    ref[0, :, 0, 0, 0] = [1.0, 1.0]
    S = 4.0 + 1e-30
    want = -(0.5 * math.log(1e-30 / S) + 0.5 * math.log(4.0 / S))
    assert _energy(oracle, c3raw, 1, ref, 1) == pytest.approx(want, rel=1e-15)


def test_energy_gibbs(oracle):
    """With q the normalised c3 of the same structure, E = sum tF H(q) (minimum);
    any other structure scores higher."""
    L, LMAX, N, K = 6, 8, 4, 3
    c = random_real_coeff(K, L, LMAX, 2, c00=6.0)
    b = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(b.c3_factored(c), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    Emin = 0.0
    for i, (k1, k2, k3) in enumerate(ktriples(K)):
        q = p[:, i] / p[:, i].sum()
        Emin += -oracle.tripletFactor(k1, k2, k3) * (q * np.log(q)).sum()
    assert _energy(oracle, p, N, ref, K) == pytest.approx(Emin, rel=1e-13)
    c2 = random_real_coeff(K, L, LMAX, 3, c00=6.0)
    assert _energy(oracle, b.c3_factored(c2), N, ref, K) > Emin


def test_energy_k3_range_walk(oracle):
    """energy() pairs column i with the i-th triple of the K3_range walk (correlations.jl:211-216)."""
    N, K = 2, 4
    rng = np.random.default_rng(0)
    ref = rng.uniform(0.1, 1.0, size=(N, 2 * N, K, K, K))
    kts = ktriples(K, lo=2)
    c3raw = rng.uniform(0.1, 1.0, size=(2 * N * N, len(kts)))
    want = 0.0
    for i, (k1, k2, k3) in enumerate(kts):
        p = c3raw[:, i] / c3raw[:, i].sum()
        q = ref[:, :, k3 - 1, k2 - 1, k1 - 1].T.ravel()
        q = q / q.sum()
        want += -(q * np.log(p)).sum() * oracle.tripletFactor(k1, k2, k3)
    assert oracle.energy(c3raw, N, ref, K, 2, K) == pytest.approx(want, rel=1e-13)


# ---------------------------------------------------------------- moves
def test_frame_rotation_equals_haar_step(oracle):
    rng = np.random.default_rng(4)
    for n in (5, 9, 37):
        C, s = oracle.haar_from_gaussian(rng.normal(size=(n, n)))
        Q, R = np.linalg.qr(rng.normal(size=(n, n)))
        assert np.abs(C @ C.T - np.eye(n)).max() < 1e-13
        B1 = oracle.rotation_step_matrix(C, 0.37)
        B2 = oracle.frame_rotation_matrix(C[0], C[1], 0.37)
        assert np.abs(B1 - B2).max() < 1e-13
        assert np.linalg.det(B1) == pytest.approx(1.0, abs=1e-12)
        assert np.trace(B1) == pytest.approx(n - 2 + 2 * math.cos(0.37), abs=1e-12)


def test_haar_matches_sign_fixed_qr(oracle):
    rng = np.random.default_rng(8)
    A = rng.normal(size=(7, 7))
    C, s = oracle.haar_from_gaussian(A)
    Q, R = np.linalg.qr(A)
    assert np.abs(C - Q * np.sign(np.diag(R))).max() < 1e-13
    assert s == (1 if np.linalg.det(C) > 0 else -1)


def test_proposal_preserves_two_photon_invariants(oracle):
    K, L, LMAX = 4, 8, 10
    c = random_real_coeff(K, L, LMAX, 6)
    st = np.array([0.3 * (l - 1) for l in range(2, L + 1, 2)])
    out = oracle.mc_propose(123, 5, oracle.TAG_MOVE, L, st, c, LMAX)
    for l in range(0, L + 1, 2):
        idx = [seanindex(l, m, LMAX) - 1 for m in range(-l, l + 1)]
        g0 = c[:, idx].conj() @ c[:, idx].T
        g1 = out[:, idx].conj() @ out[:, idx].T
        assert np.abs(g1 - g0).max() < 1e-12 * np.abs(g0).max()
        # the rotated function stays real: c_{l,-m} = (-1)^m conj(c_{l,m})
        for m in range(1, l + 1):
            a = out[:, seanindex(l, m, LMAX) - 1]
            bm = out[:, seanindex(l, -m, LMAX) - 1]
            assert np.abs(bm - (-1) ** m * a.conj()).max() < 1e-12
    assert np.abs(out - c).max() > 1e-3


def test_proposal_matches_matrix_form(oracle):
    """The two-frame proposal equals rtc R ctr applied with the explicit matrix."""
    K, L, LMAX = 2, 4, 6
    c = random_real_coeff(K, L, LMAX, 11)
    st = np.array([0.5, 0.9])
    seed, step = 77, 3
    out = oracle.mc_propose(seed, step, oracle.TAG_MOVE, L, st, c, LMAX)
    ref = c.copy()
    for li, l in enumerate(range(2, L + 1, 2)):
        n = 2 * l + 1
        g = np.array([oracle.lib().or_normal(seed, step, li, j, oracle.TAG_MOVE) for j in range(2 * n + 1)])
        u = g[:n] / np.linalg.norm(g[:n])
        v = g[n:2 * n] - (g[n:2 * n] @ u) * u
        v /= np.linalg.norm(v)
        Rm = oracle.frame_rotation_matrix(u, v, st[li] * g[2 * n])
        ref = oracle.apply_real_rotation(l, Rm, ref, LMAX)
    assert np.abs(out - ref).max() < 1e-13


def test_umat_unitary_and_real(oracle):
    for l in (0, 2, 5, 18):
        U = oracle.umat(l)
        assert np.abs(U @ U.conj().T - np.eye(2 * l + 1)).max() < 1e-15
    c = random_real_coeff(1, 6, 8, 1)
    idx = [seanindex(6, m, 8) - 1 for m in range(-6, 7)]
    y = oracle.umat(6) @ c[0, idx]
    assert np.abs(y.imag).max() < 1e-14


# ---------------------------------------------------------------- Wigner D
def test_wigner_d_orthogonality(oracle):
    """tests/test_spherical_harmonics.jl:6-27 with exact quadrature:
    int D_{m s} conj(D_{m s}) sin(theta) = 8 pi^2 / (2l+1)."""
    l = 4
    gx, gw = np.polynomial.legendre.leggauss(2 * l + 2)
    n = 2 * l + 1
    acc = np.zeros((n, n))
    for x, w in zip(gx, gw):
        D = oracle.rotation_matrix(math.acos(x), 0.3, 1.1, l)
        acc += w * np.abs(D) ** 2
    acc *= 4 * math.pi ** 2
    assert np.allclose(acc, 8 * math.pi ** 2 / (2 * l + 1), rtol=1e-12)


def test_wigner_d_group_property(oracle):
    for l in (2, 5, 8):
        D1 = oracle.rotation_matrix(0.4, 0.0, 0.0, l)
        D2 = oracle.rotation_matrix(0.9, 0.0, 0.0, l)
        D12 = oracle.rotation_matrix(1.3, 0.0, 0.0, l)
        assert np.abs(D1 @ D2 - D12).max() < 1e-13
        D = oracle.rotation_matrix(0.7, 1.2, -0.3, l)
        F = oracle.rotation_matrix(0.0, 1.2, 0.0, l) @ oracle.rotation_matrix(0.7, 0.0, 0.0, l) @ \
            oracle.rotation_matrix(0.0, 0.0, -0.3, l)
        assert np.abs(D - F).max() < 1e-13
        assert np.abs(D @ D.conj().T - np.eye(2 * l + 1)).max() < 1e-13


# ---------------------------------------------------------------- synthetic intensity
def test_intensity_two_atoms_plane_wave_expansion(oracle):
    """I = |1 + e^{i q.d}|^2 = 2 + 2 cos(q.d); with e^{i q.d} = 4pi sum_l i^l j_l(qd)
    sum_m conj(Y_lm(d)) Y_lm(q): c_lm = 8 pi i^l j_l(q|d|) conj(Y_lm(d)) (even l > 0),
    c_00 = 2 sqrt(4pi) + 8 pi j_0(q|d|) Y_00."""
    from scipy.special import spherical_jn
    atoms = np.array([[0.0, 0.0, 0.0], [0.8, -0.5, 1.1]])
    d = atoms[1]
    r = np.linalg.norm(d)
    thd, phd = math.acos(d[2] / r), math.atan2(d[1], d[0])
    KMAX, LMAX, dq = 3, 10, 0.9
    c = oracle.intensity_from_atoms(atoms, KMAX, dq, LMAX, 8, 24, 48)
    for k in range(1, KMAX + 1):
        q = k * dq
        for l in range(0, 9):
            for m in range(-l, l + 1):
                if l % 2:
                    want = 0.0
                else:
                    want = 8 * math.pi * (1j ** l) * spherical_jn(l, q * r) * np.conj(_Y(oracle, l, m, thd, phd))
                    if l == 0:
                        want += 2 * math.sqrt(4 * math.pi)
                assert abs(c[k - 1, seanindex(l, m, LMAX) - 1] - want) < 1e-11


# ---------------------------------------------------------------- annealing driver
def test_anneal_oracle_semantics(oracle):
    L, LMAX, N, K = 4, 6, 4, 3
    c = random_real_coeff(K + 1, L, LMAX, 12, c00=5.0)
    b = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(b.c3_factored(random_real_coeff(K, L, LMAX, 13, c00=5.0)), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    r = oracle.anneal(b, ref, c, 40, math.pi / 4, 1.01, 0.1, 0.99998, seed=5, ntemp=10)
    acc = r["accepted"].astype(bool)
    assert acc.any() and (~acc).any()
    # E follows E_new exactly on accepted steps and stays otherwise
    E_prev = r["E0"]
    for i in range(40):
        assert r["E"][i] == (r["E_new"][i] if acc[i] else E_prev)
        E_prev = r["E"][i]
    assert r["T_final"] == pytest.approx(r["T0"] * 0.99998 ** 40, rel=1e-12)
    # the final structure's energy is the final E
    c3 = b.c3_factored(r["coeff"])
    assert oracle.energy(c3, N, ref, K, 1, K) == pytest.approx(r["E_final"], rel=1e-13)


# ---------------------------------------------------------------- energy gradient
def _grad_case(oracle, seed=5):
    L, LMAX, N, K = 6, 8, 4, 3
    c = random_real_coeff(K, L, LMAX, seed, c00=6.0)
    b = oracle.Basis(L, LMAX, N, K)
    other = random_real_coeff(K, L, LMAX, seed + 1, c00=6.0)
    p = np.maximum(b.c3_factored(other), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    return b, c, ref, N, K


def test_energy_grad_directional_fd(oracle):
    """Re sum conj(gamma) . delta equals the central difference of the energy
    along delta (every coefficient entry an independent complex variable)."""
    b, c, ref, N, K = _grad_case(oracle)
    E, g = b.energy_grad(c, ref)
    assert E == pytest.approx(_energy(oracle, b.c3_factored(c), N, ref, K), rel=1e-14)
    rng = np.random.default_rng(9)
    for _ in range(3):
        d = rng.normal(size=c.shape) + 1j * rng.normal(size=c.shape)
        h = 1e-5
        Ep = _energy(oracle, b.c3_factored(c + h * d), N, ref, K)
        Em = _energy(oracle, b.c3_factored(c - h * d), N, ref, K)
        fd = (Ep - Em) / (2 * h)
        an = float(np.real(np.vdot(g, d)))
        assert an == pytest.approx(fd, rel=1e-6, abs=1e-9)


def test_energy_grad_single_entries(oracle):
    """Real and imaginary parts of single entries of each slot's l (the gradient
    is not symmetric in c_{l,m} and c_{l,-m}: both enter the triple products)."""
    b, c, ref, N, K = _grad_case(oracle, seed=11)
    _, g = b.energy_grad(c, ref)
    h = 1e-5
    for (k, l, m) in [(0, 2, 1), (1, 4, -3), (2, 6, 6), (2, 0, 0)]:
        i = seanindex(l, m, 8) - 1
        for unit in (1.0, 1j):
            d = np.zeros_like(c)
            d[k, i] = unit
            fd = (_energy(oracle, b.c3_factored(c + h * d), N, ref, K) -
                  _energy(oracle, b.c3_factored(c - h * d), N, ref, K)) / (2 * h)
            an = g[k, i].real if unit == 1.0 else g[k, i].imag
            assert an == pytest.approx(fd, rel=1e-5, abs=1e-8)


def test_energy_grad_zero_at_gibbs_minimum(oracle):
    """With the histogram equal to the structure's own normalised c3 the energy
    is at its minimum over positive c3, so the gradient vanishes."""
    L, LMAX, N, K = 6, 8, 4, 3
    c = random_real_coeff(K, L, LMAX, 2, c00=6.0)
    b = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(b.c3_factored(c), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    _, g = b.energy_grad(c, ref)
    _, g2 = b.energy_grad(c, _grad_case(oracle)[2])
    assert np.abs(g).max() <= 1e-12 * np.abs(g2).max()


def test_energy_grad_orthogonal_to_so3(oracle):
    """A common SO(3) rotation of all shells leaves c3 and E unchanged, so the
    gradient has no component along it."""
    b, c, ref, N, K = _grad_case(oracle, seed=21)
    _, g = b.energy_grad(c, ref)
    eps = 1e-5

    def rotated(e):
        rc = c.copy()
        for l in range(0, 7, 2):
            D = oracle.rotation_matrix(0.3 * e, 0.7 * e, -0.2 * e, l)
            idx = [seanindex(l, m, 8) - 1 for m in range(-l, l + 1)]
            rc[:, idx] = c[:, idx] @ D.T
        return rc

    d = (rotated(eps) - rotated(-eps)) / (2 * eps)   # the generator applied to c, to O(eps^2)
    assert np.linalg.norm(d) > 1.0
    assert abs(np.real(np.vdot(g, d))) <= 1e-8 * np.linalg.norm(g) * np.linalg.norm(d)
