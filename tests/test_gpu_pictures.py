"""Synthetic pictures on the device (tp_pictures) against the oracle
(oracle/picture_oracle.py, pinned in test_oracle_pictures.py) on the same
random streams, and the fused histogram generation (tp_hist_add_generated)
against tp_hist_add of the same pictures.

Bar: identical accepted-photon counts and positions within 1e-13 (the normals
go through the device's and the host's log / cos); histogram counts equal."""
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
def pic(oracle):
    from oracle import picture_oracle
    return picture_oracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _volume(tp, LMAX, K, rmax, seed):
    c = random_real_coeff(K, LMAX - 1, LMAX, seed, even=True, c00=6.0, decay=0.2)
    return tp.SphericalHarmonicsVolume(c, LMAX, K, rmax)


@pytest.mark.parametrize("LMAX,K,lam,bw", [(6, 10, 1.0, 0.0), (9, 8, 0.0, 0.05)])
def test_pictures_match_oracle(tp, pic, LMAX, K, lam, bw):
    from tp_amd import pictures
    say(f"case LMAX={LMAX} K={K} lambda={lam} beamstop={bw}")
    qcut = 1.0
    vol = _volume(tp, LMAX, K, qcut, 3 + K)
    got, rot = pictures.pointsPerOrientation(vol, qcut, 6, photons_per_image=400, lam=lam, beamstop_width=bw,
                                             capacity=400, seed=21, first_picture=5)
    surf = pic.surfaces(vol.coeff, LMAX)
    for i in range(6):
        want, R = pic.points_per_orientation(surf, LMAX, K, qcut, qcut, qcut / 3.0, 400, lam, bw, 21, 5 + i)
        say(f"  picture {5 + i}: {len(got[i])} photons (oracle {len(want)})")
        np.testing.assert_allclose(rot[i], R, atol=1e-14)
        assert len(got[i]) == len(want)
        if len(want):
            assert np.abs(got[i] - want).max() <= 1e-13


def test_generated_histograms_equal_histograms_of_the_pictures(tp):
    from tp_amd import pictures
    LMAX, K, qcut, N, K2, K3 = 8, 12, 1.0, 8, 12, 8
    vol = _volume(tp, LMAX, K, qcut, 41)
    npic = 300
    pics, _ = pictures.pointsPerOrientation(vol, qcut, npic, photons_per_image=600, capacity=512, seed=5)
    h1 = tp.PhotonHistogram(N, K2, K3, qcut / K2)
    off = np.concatenate([[0], np.cumsum([len(p) for p in pics])]).astype(np.int64)
    h1.add(np.vstack(pics), off)
    h2 = tp.PhotonHistogram(N, K2, K3, qcut / K2)
    over = pictures.generate_histograms(h2, vol, qcut, npic, photons_per_image=600, capacity=512, seed=5)
    a, b = h1.read(), h2.read()
    say(f"  {off[-1]} photons, {int(a['n3'].sum())} triplets; overflow {over}")
    assert over == 0 and a["n3"].sum() > 0
    for key in ("n1", "n2", "n3"):
        assert np.array_equal(a[key], b[key])


def test_capacity_overflow_is_reported(tp):
    from tp_amd import pictures
    vol = _volume(tp, 6, 8, 1.0, 9)
    with pytest.raises(RuntimeError, match="capacity"):
        pictures.pointsPerOrientation(vol, 1.0, 3, photons_per_image=800, capacity=2, seed=1)
    h = tp.PhotonHistogram(8, 8, 6, 1.0 / 8)
    with pytest.raises(RuntimeError, match="capacity"):
        pictures.generate_histograms(h, vol, 1.0, 3, photons_per_image=800, capacity=2, seed=1)
    # the overflowed pictures are left out whole: no photon, pair or triplet of them is counted
    a = h.read()
    assert a["n1"].sum() == 0 and a["n2"].sum() == 0 and a["n3"].sum() == 0


def test_particles_and_noise_match_oracle(tp, pic):
    """Two particles per picture, each followed by Gaussian noise photons
    (generateHistogram's segments), against the oracle on the same streams."""
    from tp_amd import pictures
    LMAX, K, qcut = 6, 10, 1.0
    vol = _volume(tp, LMAX, K, qcut, 77)
    kw = dict(photons_per_image=150, lam=0.0, capacity=512, seed=9, particles=2, noise_photons=100,
              noise_sigma=0.4)
    got, _ = pictures.pointsPerOrientation(vol, qcut, 4, **kw)
    surf = pic.surfaces(vol.coeff, LMAX)
    nsurf = pic.surfaces(pic.noise_coefficients(K, LMAX, qcut, 0.4), LMAX)
    for i in range(4):
        want = pic.picture(surf, nsurf, LMAX, K, qcut, qcut, qcut / 3.0, 150, 0.0, 0.0, 9, i, 2, 100, 0.4)
        say(f"  picture {i}: {len(got[i])} photons (oracle {len(want)})")
        assert len(got[i]) == len(want) and len(want) > 0
        assert np.abs(got[i] - want).max() <= 1e-13

    # the fused generation draws the same segments
    h1 = tp.PhotonHistogram(8, 10, 8, qcut / 10)
    off = np.concatenate([[0], np.cumsum([len(p) for p in got])]).astype(np.int64)
    h1.add(np.vstack(got), off)
    h2 = tp.PhotonHistogram(8, 10, 8, qcut / 10)
    assert pictures.generate_histograms(h2, vol, qcut, 4, **kw) == 0
    a, b = h1.read(), h2.read()
    for key in ("n1", "n2", "n3"):
        assert np.array_equal(a[key], b[key])


def test_photon_variance_and_beamstop_shots_match_oracle(tp, pic):
    """incident_photon_variance (per-particle counts) and beamstop_only
    pictures against the oracle on the same streams."""
    from tp_amd import pictures
    LMAX, K, qcut = 6, 10, 1.0
    vol = _volume(tp, LMAX, K, qcut, 78)
    got, _ = pictures.pointsPerOrientation(vol, qcut, 5, photons_per_image=200, lam=0.5, capacity=512, seed=13,
                                           first_picture=3, particles=2, incident_photon_variance=40)
    surf = pic.surfaces(vol.coeff, LMAX)
    for i in range(5):
        want = pic.picture(surf, None, LMAX, K, qcut, qcut, qcut / 3.0, 200, 0.5, 0.0, 13, 3 + i, 2, 0, 1.0,
                           variance=40)
        say(f"  picture {3 + i}: {len(got[i])} photons (oracle {len(want)})")
        assert len(got[i]) == len(want) and len(want) > 0
        assert np.abs(got[i] - want).max() <= 1e-13
    shots, rot = pictures.pointsPerOrientation(vol, qcut, 3, photons_per_image=300, capacity=512, seed=2,
                                               beamstop_width=0.1, beamstop_only=True)
    assert np.array_equal(rot, np.broadcast_to(np.eye(3), (3, 3, 3)))
    for i in range(3):
        want = pic.beamstop_shot(qcut, 300, 0.1, 2, i)
        assert len(shots[i]) == len(want) > 0
        assert np.abs(shots[i] - want).max() <= 1e-15


def test_empty_pictures(tp):
    """No incident photons, or no pictures: empty results, no error."""
    from tp_amd import pictures
    vol = _volume(tp, 6, 8, 1.0, 3)
    got, rot = pictures.pointsPerOrientation(vol, 1.0, 3, photons_per_image=0, seed=2)
    assert all(len(p) == 0 for p in got) and rot.shape == (3, 3, 3)
    got, _ = pictures.pointsPerOrientation(vol, 1.0, 0, photons_per_image=10, seed=2)
    assert got == []
    h = tp.PhotonHistogram(8, 8, 6, 1.0 / 8)
    assert pictures.generate_histograms(h, vol, 1.0, 4, photons_per_image=0, seed=2) == 0
    assert h.read()["n1"].sum() == 0
