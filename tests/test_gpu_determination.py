"""The determination driver on the device: loadHistograms' clamp ahead of the
energy (datagen.jl:432-475), the resumable annealing with the stopping rule of
rotate_all_at_once (determination.jl:1, 281-338) against the oracle's driver,
the checkpoint resume (saveState, :264-278) and rotation_search (:154-229)."""
import math
import os
import sys

import numpy as np
import pytest

from helpers import c3ref_from_columns, random_real_coeff, rel_err

pytestmark = pytest.mark.gpu
E_TOL = 1e-10


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _problem(oracle, L=4, LMAX=6, N=4, K=3, seed=12):
    c = random_real_coeff(K + 1, L, LMAX, seed, c00=5.0)
    ob = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(ob.c3_factored(random_real_coeff(K, L, LMAX, seed + 1, c00=5.0)), 1e-30)
    return ob, c3ref_from_columns(p / p.sum(), N, K), c


def test_zero_slice_is_refused_and_loadhistograms_makes_it_finite(tp, oracle):
    """A raw histogram with an all-zero k-triple slice has no normalised form:
    tp_ref_create refuses it; after loadHistograms' clamp the energy is finite
    and equals the oracle's on the loaded histogram."""
    from tp_amd import determination as det
    L, LMAX, N, K = 4, 6, 4, 4
    say("zero slice")
    c = random_real_coeff(K, L, LMAX, 31, c00=4.0)
    ob = oracle.Basis(L, LMAX, N, K)
    rng = np.random.default_rng(2)
    raw = c3ref_from_columns(rng.poisson(5.0, size=(2 * N * N, ob.klength)).astype(float), N, K)
    raw[:, :, 1, 2, 3] = 0.0     # (k1, k2, k3) = (4, 3, 2): empty
    raw[0, 0, 0, 0, 0] = 0.0     # and one empty bin elsewhere
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    with pytest.raises(tp.TPError, match="sums to 0"):
        tp.energy(vol, b, raw)
    loaded = det.load_c3(raw, K)
    E = tp.energy(vol, b, loaded)
    Eo = oracle.energy(ob.c3_factored(c), N, oracle.load_c3(raw, K), K, 1, K)
    assert np.isfinite(E) and abs(E - Eo) / abs(Eo) < E_TOL
    # an old N x N file is doubled first
    half = raw[:, :N]
    E2 = tp.energy(vol, b, det.load_c3(half, K))
    E2o = oracle.energy(ob.c3_factored(c), N, oracle.load_c3(half, K), K, 1, K)
    assert abs(E2 - E2o) / abs(E2o) < E_TOL


def test_histogram_cache_sees_in_place_changes(tp):
    L, LMAX, N, K = 4, 6, 4, 3
    c = random_real_coeff(K, L, LMAX, 3, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    ref = np.random.default_rng(4).uniform(0.5, 1.5, size=(N, 2 * N, K, K, K))
    E1 = tp.energy(vol, b, ref)
    ref[:, :, 0, 0, 0] *= 3.0
    ref[0, 0, 0, 0, 0] = 9.0
    E2 = tp.energy(vol, b, ref)
    fresh = tp.energy(vol, b, ref.copy())
    assert E2 == fresh and E2 != E1


def test_resume_from_checkpoint_is_bit_identical(tp, oracle, tmp_path):
    """300 steps in one go == 130 steps, checkpoint, a new device state from the
    file, 170 steps: the same traces and the same final structure, bit for bit."""
    from tp_amd import determination as det
    L, LMAX, N, K = 6, 8, 6, 4
    ob, ref, c = _problem(oracle, L, LMAX, N, K, 40)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K + 1, 1.0)
    say("resume: continuous")
    full = det.AnnealRun(tp.DeviceState(vol, b), b, ref, seed=9, ntemp=12).begin()
    tr_full = full.steps(300)
    say("resume: interrupted")
    part = det.AnnealRun(tp.DeviceState(vol, b), b, ref, seed=9, ntemp=12).begin()
    tr_a = part.steps(130)
    path = str(tmp_path / "state.ckpt")
    part.save(path)
    del part
    back = det.AnnealRun.load(path, b, ref)
    assert back.i == 130
    tr_b = back.steps(170)
    for key in ("E", "E_new", "accepted"):
        assert np.array_equal(np.concatenate([tr_a[key], tr_b[key]]), tr_full[key]), key
    assert np.array_equal(back.coefficients(), full.coefficients())
    assert back.E == full.E and back.T == full.T and back.stepsizes == full.stepsizes
    assert back.naccepted == full.naccepted == int(tr_full["accepted"].sum())


@pytest.mark.parametrize("min_stepsize", [2.0, 0.3])
def test_stopping_rule_matches_oracle(tp, oracle, min_stepsize):
    """Steps until stepsize_L <= min (determination.jl:322) on the device and in
    the oracle's driver: the same number of steps and the same decisions.
    min 0.3 needs several hundred steps, past the host's flag checks."""
    from tp_amd import determination as det
    ob, ref, c = _problem(oracle)
    L, LMAX, N, K = ob.L, ob.LMAX, ob.N, ob.K
    args = (math.pi / 4, 1.01, 0.1, 0.99998, 5)
    say("stopping rule", min_stepsize)
    orun = oracle.AnnealState(ob, ref, c, *args, ntemp=10)
    otr = orun.steps(20000, min_stepsize)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    st = tp.DeviceState(tp.SphericalHarmonicsVolume(c, LMAX, K + 1, 1.0), b)
    res = det.rotate_all_at_once(st, b, ref, None, *args[:4], seed=5, ntemp=10, min_stepsize=min_stepsize,
                                 checkpoint_every=20000)
    say("  steps", res["steps"], "oracle", orun.step)
    assert res["stopped"] and res["steps"] == orun.step == len(otr["E"])
    assert np.array_equal(res["accepted"], otr["accepted"])
    assert rel_err(res["E_new"], otr["E_new"]) < E_TOL
    assert res["run"].stepsizes[L] <= min_stepsize
    assert np.abs(st.read() - orun.coeff).max() < 1e-10 * np.abs(orun.coeff).max()


def _search_inputs(tp, tmp_path):
    """A histogram file made on the device: c2 and c3 of a point-atom structure,
    the triplets as Poisson counts of the normalised c3."""
    from tp_amd import determination as det
    from tp_amd import workload
    L, LMAX, N, K2, K3 = 4, 8, 8, 10, 4
    atoms = workload.atoms_in_ball(60, 6.0, 3)
    vol = tp.intensity_from_atoms(atoms, K2, math.pi, LMAX, L, 24, 48)
    b = tp.HIP_calculate_basis(L, LMAX, N, K3)
    c2 = tp.twoPhotons(vol, b, K2)
    c3raw = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    p = np.maximum(c3raw, 1e-30)
    counts = np.random.default_rng(8).poisson(1e6 * p / p.sum()).astype(float)
    c3 = c3ref_from_columns(counts, N, K3)
    path = str(tmp_path / "hist.npz")
    det.write_histograms(path, {"num_pictures": 0}, np.maximum(c2, 0.0), c3, np.ones(K2))
    params = {"histograms": path, "L": L, "LMAX": LMAX, "N": N, "K2_range": list(range(1, K2 + 1)),
              "K3_range": list(range(1, K3 + 1)), "qmax": math.pi, "lambda": 0.0, "initial_stepsize": math.pi / 4,
              "stepsizefactor": 1.01, "initial_temperature_factor": 0.1, "temperature_decay": 0.995,
              "measure": "Bayes", "seed": 17}
    return params


def test_rotation_search_resumes_to_the_same_result(tp, tmp_path):
    """rotation_search stopped half way and called again ends where an
    uninterrupted search ends: same step count, energy and structure (the
    stopping rule at 1.5 degrees; temperature decay 0.995 so that it is
    reached in a test's time)."""
    from tp_amd import determination as det
    params = _search_inputs(tp, tmp_path)
    cap = 20000   # decay 0.995 stops such a run after ~1-2 thousand steps; the cap keeps a failure finite
    say("rotation_search: uninterrupted")
    _, s1 = det.rotation_search(params, str(tmp_path / "a"), nsteps=cap)
    assert s1["state"] == "finished_structure"
    say("  steps", s1["i"], "E", s1["E"])
    assert s1["stepsizes"][str(params["L"])] <= det.minstepsize
    cut = s1["i"] // 2 + 1
    say("rotation_search: interrupted at", cut)
    _, s2 = det.rotation_search(params, str(tmp_path / "b"), nsteps=cut)
    assert s2["state"] == "running" and s2["i"] == cut
    info = det.read_checkpoint_info(str(tmp_path / "b" / "state.ckpt"))
    assert info.status.step == cut
    _, s3 = det.rotation_search(params, str(tmp_path / "b"), nsteps=cap - cut)
    assert s3["state"] == "finished_structure"
    assert s3["i"] == s1["i"] and s3["E"] == s1["E"]
    c1 = s1["result"].coeff if hasattr(s1["result"], "coeff") else s1["result"]
    c3 = s3["result"].coeff if hasattr(s3["result"], "coeff") else s3["result"]
    assert np.array_equal(c1, c3)
    # a finished search is not started again
    _, s4 = det.rotation_search(params, str(tmp_path / "b"))
    assert s4["state"] == "finished_structure" and "traces" not in s4
    log = open(str(tmp_path / "a" / "det.out")).read().splitlines()
    assert sum(1 for x in log if not x.startswith("#")) == s1["i"]
