"""The device's photon-pair / photon-triplet histograms (tp_hist_*) against
the oracle's restatement of histogramCorrelationsInPicture_alltoall
(or_hist_pictures, pinned in test_oracle_hist.py).  Bar: integer counts
identical (bin indexing exact); weighted histograms within 1e-12 of the
oracle's sequential sums."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.mark.parametrize("N,K2,K3,npic,nphot,bs", [
    (4, 6, 4, 50, 12, 0.0),
    (8, 12, 9, 40, 60, 0.05),
    (32, 38, 26, 12, 200, 0.02),    # the benchmark's binning, ~150 photons per picture
])
def test_hist_matches_oracle(tp, oracle, N, K2, K3, npic, nphot, bs):
    from tp_amd import workload
    say(f"case N={N} K2={K2} K3={K3} pictures={npic} photons={nphot}")
    qcut = 1.0
    dq = qcut / K2
    ph, off = workload.beamstop_pictures(npic, nphot, qcut, seed=7 + N, beamstop_width=bs)
    h = tp.PhotonHistogram(N, K2, K3, dq)
    h.add(ph, off)
    got = h.read()
    c1, c2, c3, n1, n2, n3 = oracle.hist_pictures(ph, off, dq, N, K2, K3)
    assert np.array_equal(got["n1"], n1) and np.array_equal(got["n2"], n2) and np.array_equal(got["n3"], n3)
    assert n3.sum() > 0
    for a, b in ((got["c1"], c1), (got["c2"], c2), (got["c3"], c3)):
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(b))
    say(f"  counts: {int(n1.sum())} photons, {int(n2.sum())} pairs, {int(n3.sum())} triplets")


def test_hist_split_calls_identical(tp):
    """Histograms are exact integer sums: two calls over halves equal one call."""
    from tp_amd import workload
    N, K2, K3 = 8, 12, 9
    ph, off = workload.beamstop_pictures(30, 40, 1.0, seed=3)
    h1 = tp.PhotonHistogram(N, K2, K3, 1.0 / K2)
    h1.add(ph, off)
    h2 = tp.PhotonHistogram(N, K2, K3, 1.0 / K2)
    h2.add(ph[:off[13]], off[:14])
    h2.add(ph[off[13]:], off[13:] - off[13])
    a, b = h1.read(), h2.read()
    for k in ("n1", "n2", "n3", "c1", "c2", "c3"):
        assert np.array_equal(a[k], b[k])


def test_hist_rejects_oversized_picture(tp):
    h = tp.PhotonHistogram(4, 6, 4, 1.0 / 6)
    ph = np.zeros((600, 3))
    ph[:, 0] = 0.5
    with pytest.raises(tp.TPError):
        h.add(ph, np.array([0, 600]))
