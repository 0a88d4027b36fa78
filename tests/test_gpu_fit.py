"""Rotational fit on the device (tp_fit_*) against the oracle (oracle/fit_oracle.py,
pinned in test_oracle_fit.py): the similarity of rotated copies, and the two
searches built on it.

Bar: |gpu - oracle| <= 1e-10 for similarities (closed-form phi sums and
quaternion Wigner-D against point-wise surfaces and the factorial d_work)."""
import sys

import numpy as np
import pytest

from helpers import random_real_coeff

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


@pytest.fixture(scope="module")
def fo(oracle):
    from oracle import fit_oracle
    return fit_oracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("L,LMAX,K,klo,even", [
    (8, 10, 5, 2, False),     # odd l and l > L present: kept unrotated
    (4, 6, 3, 1, True),
    (18, 25, 12, 3, True),    # the full structure's L and LMAX
])
def test_similarity_matches_oracle(tp, fo, L, LMAX, K, klo, even):
    from tp_amd.fit import StructureFit
    say(f"case L={L} LMAX={LMAX} K={K} klo={klo} even={even}")
    lfill = L if even else LMAX - 1
    vol = random_real_coeff(K, lfill, LMAX, 100 + L, even=even, c00=3.0, decay=0.05)
    ref = random_real_coeff(K, lfill, LMAX, 200 + L, even=even, c00=3.0, decay=0.05)
    b = tp.HIP_calculate_basis(L, LMAX, 2, 1)
    rng = np.random.default_rng(L)
    euler = np.vstack([[0.0, 0.0, 0.0], rng.uniform(-3.5, 3.5, size=(5 if L < 18 else 2, 3))])
    fit = StructureFit(b, ref, range(klo, K + 1)).set_volume(vol)
    got = fit.similarity(euler)
    for i, (t, p, g) in enumerate(euler):
        want = fo.similarity(ref, fo.rotate_structure(vol, t, p, g, K, range(2, L + 1, 2), LMAX), LMAX, klo, K)
        say(f"  euler {np.round(euler[i], 3)}: gpu {got[i]:.15f} oracle {want:.15f}")
        assert abs(got[i] - want) <= 1e-10
    # the structure with itself
    self_fit = StructureFit(b, vol, range(klo, K + 1)).set_volume(vol)
    assert self_fit.similarity([[0.0, 0.0, 0.0]])[0] == pytest.approx(1.0, abs=1e-12)


def test_rotate_structure_matches_oracle(tp, fo):
    from tp_amd.fit import rotateStructure
    L, LMAX, K = 8, 10, 4
    c = random_real_coeff(K + 2, LMAX - 1, LMAX, 7, even=False)
    b = tp.HIP_calculate_basis(L, LMAX, 2, 1)
    got = rotateStructure(c, 0.7, 1.9, -0.4, K, L, b)
    want = fo.rotate_structure(c, 0.7, 1.9, -0.4, K, range(2, L + 1, 2), LMAX)
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()


def test_fit_full_recovers_a_grid_rotation(tp, fo):
    from tp_amd.fit import fitStructures_full
    L, LMAX, K = 8, 10, 5
    ref = random_real_coeff(K, L, LMAX, 31, c00=3.0, decay=0.05)
    t, p, g = np.pi / 4, np.pi / 2, np.pi / 4          # a point of the num = 4 grid
    vol = fo.rotate_structure(ref, -t, -g, -p, K, range(2, L + 1, 2), LMAX)   # the inverse rotation
    b = tp.HIP_calculate_basis(L, LMAX, 2, 1)
    rot, sc, bt, bp, bg = fitStructures_full(vol, ref, 4, K, L, b)
    say(f"  best {sc:.15f} at {bt:.4f} {bp:.4f} {bg:.4f}")
    assert sc == pytest.approx(1.0, abs=1e-12)
    assert (bt, bp, bg) == pytest.approx((t, p, g), abs=1e-12)
    assert np.abs(rot - ref).max() <= 1e-10 * np.abs(ref).max()


def test_fit_random_finds_the_rotation(tp, fo):
    """The reference's random walk (8 repetitions, step factor 0.97) on a
    randomly rotated copy: calibrated with the oracle scorer, two repetitions
    reach > 0.999."""
    from tp_amd.fit import fitStructures_random
    L, LMAX, K = 8, 10, 5
    ref = random_real_coeff(K, L, LMAX, 21, c00=3.0, decay=0.05)
    vol = fo.rotate_structure(ref, -1.1, -0.4, -2.0, K, range(2, L + 1, 2), LMAX)
    b = tp.HIP_calculate_basis(L, LMAX, 2, 1)
    rot, sc, t, p, g = fitStructures_random(vol, ref, K, L, 0.97, 8, b, seed=1)
    say(f"  best {sc:.6f} at {t:.4f} {p:.4f} {g:.4f}")
    assert sc > 0.999
    assert fo.similarity(ref, rot, LMAX, 1, K) == pytest.approx(sc, abs=1e-10)
