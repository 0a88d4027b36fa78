"""Host logic of the rotational fit (tp_amd.fit): the Euler-to-quaternion map
against the rotation it must equal, and the random search's loop over a cheap
synthetic scorer (no library call)."""
import math

import numpy as np

from helpers import quat_to_rot


def Rz(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, -s, 0], [s, c, 0], [0, 0, 1]])


def Ry(a):
    c, s = np.cos(a), np.sin(a)
    return np.array([[c, 0, s], [0, 1, 0], [-s, 0, c]])


def test_euler_quat_is_rz_ry_rz():
    from tp_amd.fit import euler_quat
    rng = np.random.default_rng(0)
    for _ in range(20):
        t, p, g = rng.uniform(-4, 4, size=3)
        q = euler_quat(t, p, g)
        assert abs(np.linalg.norm(q) - 1) < 1e-15
        np.testing.assert_allclose(quat_to_rot(q), Rz(p) @ Ry(-t) @ Rz(g), atol=1e-14)


def test_random_search_loop():
    from tp_amd.fit import random_search
    target = np.array([0.8, 2.0, -1.0])
    calls = []

    def score(e):
        calls.append(e.copy())
        return np.cos(np.linalg.norm(e - target, axis=1) / 4)

    f, R = 0.9, 5
    best, th, ph, ga = random_search(score, R, f, np.random.default_rng(3))
    nsteps = math.ceil(math.log(1 / 180) / math.log(f))
    assert len(calls) == nsteps and all(c.shape == (R, 3) for c in calls)
    np.testing.assert_allclose(best, score(np.stack([th, ph, ga], 1)), rtol=0, atol=0)
    # best is the running maximum of each repetition's accepted scores
    seen = np.max(np.stack([np.cos(np.linalg.norm(c - target, axis=1) / 4) for c in calls[:-1]]), axis=0)
    assert np.all(best >= seen)
