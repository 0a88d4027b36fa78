"""Host code of the determination driver, on the CPU (no HIP call):
loadHistograms (tp_load_histograms, datagen.jl:432-475) against the oracle's
restatement and hand cases, the histogram file, and the annealing checkpoint
(tp_checkpoint_save / _load).  The oracle's resumable annealing driver is
pinned here against its plain driver and the stopping rule's definition."""
import ctypes as C
import math
import os

import numpy as np
import pytest

from helpers import c3ref_from_columns, random_real_coeff


def _lib():
    from tp_amd import _lib
    if not os.path.exists(_lib.LIB_PATH):
        _lib.build()
    return _lib


# ---- loadHistograms ---------------------------------------------------------------

def test_load_c3_hand_case_doubling():
    """An old N x N histogram: both beta halves carry the slice (k1 >= k2 >= k3
    only), every other entry is the clamp 1e-30, and the result sums to 1."""
    from tp_amd import determination as det
    N, K = 2, 2
    h = np.zeros((N, N, K, K, K))
    h[:, :, 0, 0, 0] = [[1.0, 2.0], [3.0, 4.0]]
    h[:, :, 0, 0, 1] = [[5.0, 0.0], [0.0, 6.0]]      # k1=2, k2=1, k3=1; two empty bins
    h[:, :, 1, 0, 1] = 7.0                          # k3 > k2: not a k1 >= k2 >= k3 slice, dropped
    out = det.load_c3(h, K)
    assert out.shape == (N, 2 * N, K, K, K)
    tot = 1 + 2 + 3 + 4 + 5 + 6
    tot = 2 * tot + 1e-30 * (out.size - 2 * 8 + 2 * 2)   # 2*4 + 2*4 filled, of which 4 are empty bins
    assert np.allclose(out[:, :N, 0, 0, 0] * tot, [[1, 2], [3, 4]], rtol=1e-14)
    assert np.array_equal(out[:, :N, 0, 0, 0], out[:, N:, 0, 0, 0])
    assert np.array_equal(out[:, :N, 0, 0, 1], out[:, N:, 0, 0, 1])
    assert out[0, 1, 0, 0, 1] == out[1, 0, 1, 0, 1] == pytest.approx(1e-30 / tot, rel=1e-14)
    assert out.sum() == pytest.approx(1.0, rel=1e-14)
    assert np.all(out > 0)


def test_load_c3_matches_oracle_and_cuts():
    """N x 2N (no doubling) and N x N inputs, cut from Kf = 5 to K3 = 3, with
    empty slices: the library equals the oracle bit for bit."""
    from oracle import pyoracle as po
    from tp_amd import determination as det
    rng = np.random.default_rng(3)
    for NB_factor in (1, 2):
        N, Kf, K3 = 4, 5, 3
        h = rng.poisson(2.0, size=(N, NB_factor * N, Kf, Kf, Kf)).astype(float)
        h[:, :, 0, 1, 2] = 0.0   # an empty slice
        got = det.load_c3(h, K3)
        want = po.load_c3(h, K3)
        assert np.array_equal(got, want)
        if NB_factor == 2:   # without doubling it is max.(c,1e-30) cut and normalised
            ref = np.maximum(h[:, :, :K3, :K3, :K3], 1e-30)
            assert np.allclose(got, ref / ref.sum(), rtol=1e-13, atol=0)
        # an empty slice becomes uniform: log q is finite everywhere
        s = got[:, :, 0, 1, 2]
        assert np.all(s == s.flat[0]) and s.flat[0] > 0


def test_load_c3_rejects_bad_shapes_and_nan():
    from tp_amd import TPError
    from tp_amd import determination as det
    with pytest.raises(TPError):
        det.load_c3(np.ones((3, 5, 2, 2, 2)), 2)      # NB neither N nor 2N
    with pytest.raises(TPError):
        det.load_c3(np.ones((3, 6, 2, 2, 2)), 3)      # K3 > Kf
    h = np.ones((2, 4, 2, 2, 2))
    h[0, 0, 0, 0, 0] = np.nan
    with pytest.raises(TPError, match="NaN"):
        det.load_c3(h, 2)


def test_histogram_file_round_trip(tmp_path):
    """write_histograms / loadHistograms: the reference's tuple (c2_full, c2,
    c3_full, c3, c1_full) with clamps, cuts and normalisations."""
    from tp_amd import determination as det
    rng = np.random.default_rng(5)
    N, K2, K3 = 4, 6, 3
    c1 = rng.poisson(3.0, size=K2).astype(float)
    c2 = rng.poisson(3.0, size=(N, K2, K2)).astype(float)
    c3 = rng.poisson(1.0, size=(N, N, K3 + 1, K3 + 1, K3 + 1)).astype(float)
    params = {"N": N, "K2": K2, "K3": K3 + 1, "num_pictures": 12}
    path = str(tmp_path / "hist.npz")
    det.write_histograms(path, params, c2, c3, c1)
    p2, c2b, c3b, c1b = det.read_histograms(path)
    assert p2 == params and np.array_equal(c2b, c2) and np.array_equal(c3b, c3) and np.array_equal(c1b, c1)
    c2_full, c2n, c3_full, c3n, c1_full = det.loadHistograms(K2 - 1, K3, path)
    assert c3_full.shape == (N, 2 * N, K3 + 1, K3 + 1, K3 + 1) and np.all(c3_full >= 1e-30)
    assert np.array_equal(c1_full, np.maximum(c1, 1e-30))
    want2 = np.maximum(c2, 1e-30)[:, :K2 - 1, :K2 - 1]
    assert np.allclose(c2n, want2 / want2.sum(), rtol=1e-14)
    from oracle import pyoracle as po
    assert np.array_equal(c3n, po.load_c3(c3, K3))


# ---- checkpoint files -------------------------------------------------------------

def _info(_lib, L=6, nshell=5, ld=64):
    st = _lib.AnnealStatus()
    st.step, st.E, st.T, st.E0, st.T0, st.nl, st.naccepted, st.stopped = 1234, 1.5, 0.25, 2.5, 0.5, L // 2, 77, 0
    for q in range(L // 2):
        st.stepsizes[q] = 0.1 * (q + 1) / 3.0
    p = _lib.AnnealParams(2000, 50, math.pi / 4, 1.01, 0.1, 0.99998, 987654321)
    return _lib.CheckpointInfo(p, st, 0.0261799, 0.0, 0.0826, L, 8, 16, 4, 1, 4, nshell, ld)


def test_checkpoint_round_trip(tmp_path):
    _l = _lib()
    L = _l.lib()
    info = _info(_l)
    rng = np.random.default_rng(0)
    coeff = rng.normal(size=(info.nshell, info.ld)) + 1j * rng.normal(size=(info.nshell, info.ld))
    path = str(tmp_path / "run.ckpt")
    _l.check(L.tp_checkpoint_save(os.fsencode(path), C.byref(info), coeff.ctypes.data), "save")
    assert not os.path.exists(path + ".tmp")
    back = _l.CheckpointInfo()
    out = np.zeros_like(coeff)
    _l.check(L.tp_checkpoint_load(os.fsencode(path), C.byref(back), out.ctypes.data, out.size * 2), "load")
    assert np.array_equal(out, coeff)   # bit for bit
    assert bytes(back) == bytes(info)
    # a buffer too small is refused
    assert L.tp_checkpoint_load(os.fsencode(path), C.byref(back), out.ctypes.data, out.size * 2 - 1) != 0


@pytest.mark.parametrize("damage", ["flip", "truncate", "magic", "missing"])
def test_checkpoint_damage_is_detected(tmp_path, damage):
    _l = _lib()
    L = _l.lib()
    info = _info(_l, nshell=2, ld=64)
    coeff = np.ones((2, 64), np.complex128)
    path = str(tmp_path / "run.ckpt")
    _l.check(L.tp_checkpoint_save(os.fsencode(path), C.byref(info), coeff.ctypes.data), "save")
    data = bytearray(open(path, "rb").read())
    if damage == "flip":
        data[200] ^= 0x10
    elif damage == "truncate":
        data = data[:-100]
    elif damage == "magic":
        data[0:6] = b"XXXXXX"
    if damage == "missing":
        os.remove(path)
    else:
        open(path, "wb").write(bytes(data))
    back = _l.CheckpointInfo()
    assert L.tp_checkpoint_load(os.fsencode(path), C.byref(back), None, 0) == -1
    assert L.tp_last_error()


# ---- the oracle's resumable driver --------------------------------------------------

def _small(po):
    L, LMAX, N, K = 4, 6, 4, 3
    c = random_real_coeff(K + 1, L, LMAX, 12, c00=5.0)
    ob = po.Basis(L, LMAX, N, K)
    p = np.maximum(ob.c3_factored(random_real_coeff(K, L, LMAX, 13, c00=5.0)), 1e-30)
    return ob, c3ref_from_columns(p / p.sum(), N, K), c


def test_oracle_resumable_driver_equals_plain_driver(oracle):
    """begin + 25 + 35 steps == or_anneal's 60 steps, bit for bit."""
    ob, ref, c = _small(oracle)
    args = (math.pi / 4, 1.01, 0.1, 0.99998, 5)
    plain = oracle.anneal(ob, ref, c, 60, *args, ntemp=10)
    run = oracle.AnnealState(ob, ref, c, *args, ntemp=10)
    assert run.E == plain["E0"] and run.T == plain["T0"]
    a, b = run.steps(25), run.steps(35)
    assert np.array_equal(np.concatenate([a["E"], b["E"]]), plain["E"])
    assert np.array_equal(np.concatenate([a["accepted"], b["accepted"]]), plain["accepted"])
    assert np.array_equal(run.coeff, plain["coeff"]) and run.step == 60 and run.T == plain["T_final"]


def test_oracle_stopping_rule(oracle):
    """The run stops before the first step at which stepsize_L <= min would
    hold, i.e. right after the step where the accepted/rejected balance first
    brings initial (L-1) f^(acc - rej) to or below min (determination.jl:322)."""
    ob, ref, c = _small(oracle)
    f, init, mn = 1.01, math.pi / 4, 2.0
    run = oracle.AnnealState(ob, ref, c, init, f, 0.1, 0.99998, 5, ntemp=10)
    tr = run.steps(5000, mn)
    acc = tr["accepted"].astype(int)
    bal = np.cumsum(2 * acc - 1)
    size = init * (ob.L - 1) * f ** bal.astype(float)
    first = int(np.argmax(size <= mn * (1 + 1e-12)))
    assert size[first] <= mn * (1 + 1e-12) and np.all(size[:first] > mn * (1 - 1e-12))
    assert len(acc) == first + 1 and run.step == first + 1
    assert run.stepsizes[-1] <= mn
    assert len(run.steps(10, mn)["E"]) == 0   # no step once the rule holds
