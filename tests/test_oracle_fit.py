"""The rotational-fit oracle (oracle/fit_oracle.py) pinned without the library:
the Euler convention of rotateStructure as a rotation of the function, and the
similarity's invariances."""
import numpy as np
import pytest

from helpers import random_real_coeff, sh_eval


@pytest.fixture(scope="module")
def fo(oracle):
    from oracle import fit_oracle
    return fit_oracle


def Rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def Ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def test_rotate_structure_is_a_rotation_of_the_function(oracle, fo):
    L, LMAX, K = 6, 8, 2
    c = random_real_coeff(K, L, LMAX, 3)
    th, ph, ga = 0.7, 1.9, -0.4
    rc = fo.rotate_structure(c, th, ph, ga, K, range(2, L + 1, 2), LMAX)
    R = Rz(ph) @ Ry(-th) @ Rz(ga)
    rng = np.random.default_rng(1)
    for _ in range(6):
        v = rng.normal(size=3)
        v /= np.linalg.norm(v)
        for l in range(2, L + 1, 2):
            for k in range(K):
                assert abs(sh_eval(oracle, rc[k], l, LMAX, v) - sh_eval(oracle, c[k], l, LMAX, R.T @ v)) < 1e-12
    assert np.array_equal(rc[:, 0], c[:, 0])


def test_similarity_invariances(fo):
    L, LMAX, K = 6, 8, 3
    c = random_real_coeff(K, L, LMAX, 5, c00=3.0)
    assert fo.similarity(c, c, LMAX, 1, K) == pytest.approx(1.0, abs=1e-14)
    aff = 2.5 * c
    aff[:, 0] += 7.0                      # a constant on every surface
    assert fo.similarity(c, aff, LMAX, 1, K) == pytest.approx(1.0, abs=1e-13)
    assert fo.similarity(c, -c, LMAX, 2, K) == pytest.approx(-1.0, abs=1e-14)
    r = fo.rotate_structure(c, 0.9, 0.3, 1.2, K, range(2, L + 1, 2), LMAX)
    s = fo.similarity(c, r, LMAX, 1, K)
    assert -1.0 < s < 0.99
