"""Two-photon correlation on the device (tp_two_photons) and the start
structure from it (tp_retrieve_solution, host code behind the same C ABI)
against the numpy oracle (oracle/c2_oracle.py, pinned in test_oracle_c2.py).

Bar: |gpu - oracle| <= 1e-12 max |oracle| for c2 (FP64, sums in the same
order up to the Legendre evaluation); retrieved coefficients within 1e-9 of
max |oracle| (Jacobi against LAPACK eigenvectors, same sign convention), the
same flip decision, and the round trip c2 -> structure -> c2 through the
device within 1e-9."""
import sys

import numpy as np
import pytest

from helpers import random_real_coeff, seanindex

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


@pytest.fixture(scope="module")
def c2o(oracle):
    from oracle import c2_oracle
    return c2_oracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def restricted_coeff(K, L, LMAX, seed, c00=4.0):
    c = random_real_coeff(K, L, LMAX, seed, c00=c00, decay=0.1)
    for k in range(1, K + 1):
        for l in range(2 * k - 1, LMAX):
            for m in range(-l, l + 1):
                c[k - 1, seanindex(l, m, LMAX) - 1] = 0.0
    return c


@pytest.mark.parametrize("L,LMAX,N,K,K2,minimal,normalize,pad", [
    (4, 6, 4, 3, 3, False, False, 0),
    (8, 10, 8, 6, 9, False, False, 0),     # more shells than the basis
    (8, 10, 8, 6, 6, True, False, 5),      # minimal, padded rows
    (6, 8, 16, 4, 12, False, True, 0),     # normalized
    (18, 19, 32, 5, 38, False, False, 0),  # histogram size, l up to 18
])
def test_two_photons_matches_oracle(tp, c2o, L, LMAX, N, K, K2, minimal, normalize, pad):
    say(f"case L={L} LMAX={LMAX} N={N} K={K} K2={K2} minimal={minimal} normalize={normalize} pad={pad}")
    c = random_real_coeff(K2, L, LMAX, 2 * L + K2, c00=3.0, decay=0.05)
    if pad:
        c = np.concatenate([c, np.full((K2, pad), 7.0 + 1j)], axis=1)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    got = tp.twoPhotons(c, b, K2, minimal, normalize)
    ref = c2o.two_photons(c[:, :LMAX * LMAX], L, LMAX, N, K2, minimal, normalize)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say(f"  rel err {err:.2e}")
    assert got.shape == (N, K2, K2)
    assert err <= 1e-12
    if minimal:
        assert all(np.all(got[:, k2, k1] == 0) for k1 in range(K2) for k2 in range(k1 + 1, K2))


@pytest.mark.parametrize("L,LMAX,N,K2,lo,c00", [
    (4, 6, 8, 10, 1, 4.0),
    (8, 10, 16, 24, 1, 4.0),
    (8, 10, 16, 24, 1, 30.0),
    (4, 6, 12, 14, 3, 4.0),
    (18, 19, 32, 38, 1, 10.0),
])
def test_retrieve_matches_oracle_and_round_trips(tp, c2o, L, LMAX, N, K2, lo, c00):
    say(f"case L={L} LMAX={LMAX} N={N} K2={K2} lo={lo} c00={c00}")
    c = restricted_coeff(K2, L, LMAX, 7 + L, c00)
    b = tp.HIP_calculate_basis(L, LMAX, N, min(K2, 6))
    c2 = tp.twoPhotons(c, b, K2)
    ro, info = c2o.retrieve_solution(c2, L, LMAX, lo, K2)
    vol, li = tp.retrieveSolution(c2, L, LMAX, range(lo, K2 + 1), 1.0, return_info=True)
    err = np.abs(vol.coeff - ro).max() / np.abs(ro).max()
    say(f"  coeff rel err {err:.2e}  flipped {li['flipped']}  negativity {li['negativity']:.6g}")
    assert vol.coeff.shape == (K2, LMAX * LMAX) and vol.KMAX == K2
    assert li["flipped"] == info["flipped"]
    assert li["negativity"] == pytest.approx(info["negativity"], rel=1e-10, abs=1e-14)
    assert err <= 1e-9
    back = tp.twoPhotons(vol, b, K2)
    s = slice(lo - 1, K2)
    rt = np.abs(back[:, s, s] - c2[:, s, s]).max() / np.abs(c2).max()
    say(f"  round trip rel err {rt:.2e}")
    assert rt <= 1e-9


def test_retrieve_rejects_bad_ranges(tp):
    c2 = np.zeros((8, 10, 10))
    with pytest.raises(RuntimeError, match="L < K2/2"):
        tp.retrieveSolution(c2, 6, 8, 10, 1.0)
    with pytest.raises(RuntimeError, match="k2hi"):
        tp.retrieveSolution(c2, 2, 8, range(1, 12), 1.0)
    with pytest.raises(RuntimeError, match="alpha_star"):
        tp.retrieveSolution(c2, 2, 8, 10, 20.0, lam=1.0)   # qmax lambda > 4 pi


@pytest.mark.parametrize("lam", [2.5, 1.0])
def test_two_photons_and_retrieve_ewald_match_oracle(tp, c2o, lam):
    """lambda != 0: twoPhotons at alpha_star with the volume's dq
    (tp_two_photons_ewald) and retrieveSolution with alpha_star columns
    (tp_retrieve_solution_ewald), against the oracle (pinned by the rotation
    average on the Ewald sphere and the round trip in test_oracle_c2.py)."""
    L, LMAX, N, K2, qmax = 4, 6, 10, 12, 1.2
    c = random_real_coeff(K2, L, LMAX, 23, c00=4.0)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K2, qmax)
    b = tp.HIP_calculate_basis(L, LMAX, N, K2, lam, vol.dq())
    got = tp.twoPhotons(vol, b, K2)
    want = c2o.two_photons(c, L, LMAX, N, K2, lam=lam, dq=vol.dq())
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    rec, info = tp.retrieveSolution(want, L, LMAX, K2, qmax, lam=lam, return_info=True)
    wrec, winfo = c2o.retrieve_solution(want, L, LMAX, 1, K2, qmax=qmax, lam=lam)
    assert info["flipped"] == winfo["flipped"]
    assert np.abs(rec.coeff - wrec).max() <= 1e-8 * np.abs(wrec).max()
    with pytest.raises(RuntimeError, match="tp_two_photons_ewald"):
        from tp_amd import check, lib
        out = np.zeros(N * K2 * K2)
        cc = np.ascontiguousarray(c)
        check(lib().tp_two_photons(b.ctx.h, b.h, cc.ctypes.data, LMAX * LMAX, K2, 0, 0, out.ctypes.data), "x")
