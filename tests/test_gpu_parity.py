"""Parity of the gfx950 path (libtp.so through its C ABI) with the CPU oracle.

Bar: c3 and the Bayes energy within 1e-10 relative of the FP64
oracle; histogram bin indexing exact.  c3 error is max|gpu - oracle| /
max|oracle| over the whole (R, klength) matrix."""
import math
import sys

import numpy as np
import pytest

from helpers import c3ref_from_columns, ktriples, random_real_coeff, rel_err, seanindex

pytestmark = pytest.mark.gpu

C3_TOL = 1e-10
E_TOL = 1e-10


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    # (L, LMAX, N, K)       tiny, ragged (R and klength not tile multiples), the small config,
    # two k1 halves of the MFMA triple products (K=20), K > 27 (the VALU triple-product kernel),
    # N = 32 (the separated-grid contraction at its largest) and N > 32 (the GEMM contraction)
    (4, 6, 4, 3),
    (6, 9, 5, 5),
    (2, 3, 3, 2),
    (8, 25, 16, 8),
    (6, 8, 4, 20),
    (4, 6, 3, 29),
    (4, 6, 32, 4),
    (4, 6, 40, 3),
]


@pytest.mark.parametrize("L,LMAX,N,K", CASES)
def test_c3_matches_oracle(tp, oracle, L, LMAX, N, K):
    say("c3 case", L, LMAX, N, K)
    c = random_real_coeff(K, L, LMAX, seed=L * 100 + N, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    ob = oracle.Basis(L, LMAX, N, K)
    assert (b.basislen, b.n_lt, b.klength, b.R) == (ob.basislen, ob.n_lt, ob.klength, ob.R)
    got = tp.FullCorrelation_parallized(tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0), b, return_raw=True)
    want = ob.c3_factored(c)
    assert rel_err(got, want) < C3_TOL
    cols = min(ob.klength, 12)
    assert rel_err(got[:, :cols], ob.c3_dense(c, 0, cols)) < C3_TOL


def test_c3_five_dim_layout(tp, oracle):
    L, LMAX, N, K = 4, 6, 4, 3
    c = random_real_coeff(K, L, LMAX, seed=3)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    t = tp.FullCorrelation_parallized(tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0), b)
    for (k1, k2, k3) in ktriples(K):
        s = oracle.c3_slice(c, N, L, LMAX, k1, k2, k3).reshape(2 * N, N).T
        assert rel_err(t[:, :, k3 - 1, k2 - 1, k1 - 1], s) < C3_TOL
    assert np.all(t[:, :, 1, 0, 0] == 0.0)   # k1 >= k2 >= k3 only


@pytest.mark.parametrize("L,LMAX,N,K", CASES)
def test_energy_matches_oracle(tp, oracle, L, LMAX, N, K):
    say("energy case", L, LMAX, N, K)
    c = random_real_coeff(K, L, LMAX, seed=7 + L, c00=4.0)
    ob = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(ob.c3_factored(random_real_coeff(K, L, LMAX, seed=99, c00=4.0)), 1e-30)
    rng = np.random.default_rng(1)
    ref = c3ref_from_columns(rng.poisson(1e4 * p / p.sum() * p.shape[1]).astype(float), N, K)
    ref = np.maximum(ref, 1e-30)
    ref /= ref.sum()
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    E = tp.energy(vol, b, ref)
    Eo = oracle.energy(ob.c3_factored(c), N, ref, K, 1, K)
    assert abs(E - Eo) / abs(Eo) < E_TOL
    if K >= 3:   # K3_range = 2:K walk (correlations.jl:211-216)
        E2 = tp.energy(vol, b, ref, range(2, K + 1))
        E2o = oracle.energy(ob.c3_factored(c), N, ref, K, 2, K)
        assert abs(E2 - E2o) / abs(E2o) < E_TOL


def test_energy_bin_indexing_exact(tp, oracle):
    """Tag every histogram entry with a distinct value: any mis-mapped bin or
    k-triple changes the energy far beyond rounding."""
    L, LMAX, N, K = 4, 6, 3, 4
    c = random_real_coeff(K, L, LMAX, seed=21, c00=5.0)
    ref = np.arange(1, N * 2 * N * K ** 3 + 1, dtype=float).reshape((N, 2 * N, K, K, K), order="F")
    ref = ref ** 1.5
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    ob = oracle.Basis(L, LMAX, N, K)
    E = tp.energy(tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0), b, ref)
    Eo = oracle.energy(ob.c3_factored(c), N, ref, K, 1, K)
    assert abs(E - Eo) / abs(Eo) < 1e-12


def test_energy_clamped_bins(tp, oracle):
    """Coefficients whose c3 has negative bins: the 1e-30 clamp (correlations.jl:203)."""
    L, LMAX, N, K = 6, 8, 6, 3
    c = random_real_coeff(K, L, LMAX, seed=4, c00=0.3)
    ob = oracle.Basis(L, LMAX, N, K)
    c3 = ob.c3_factored(c)
    assert (c3 < 0).any()
    ref = c3ref_from_columns(np.abs(c3) + 1e-3, N, K)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    E = tp.energy(tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0), b, ref)
    Eo = oracle.energy(c3, N, ref, K, 1, K)
    assert abs(E - Eo) / abs(Eo) < E_TOL


def test_state_energy_and_determinism(tp):
    L, LMAX, N, K = 8, 25, 16, 8
    c = random_real_coeff(K + 2, L, LMAX, seed=2, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    ref = np.random.default_rng(3).uniform(0.1, 1, size=(N, 2 * N, K, K, K))
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K + 2, 1.0)
    E1 = tp.energy(vol, b, ref)
    st = tp.DeviceState(vol, b)
    E2 = st.energy(ref)
    E3 = st.energy(ref)
    assert E1 == E2 == E3      # same kernels, fixed reduction order: bitwise
    back = st.read()
    assert np.array_equal(back, c)


def test_state_energy_follows_histogram_mode_and_state(tp):
    """Device-resident energies across alternating histograms, K3_ranges and
    precision modes, a new histogram after another was dropped, and a state
    moved by annealing: each equals the direct evaluation of the same
    structure (tp_energy), bit for bit."""
    L, LMAX, N, K = 8, 25, 16, 8
    c = random_real_coeff(K + 2, L, LMAX, seed=5, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    rng = np.random.default_rng(6)
    ra = rng.uniform(0.1, 1, size=(N, 2 * N, K, K, K))
    rb = rng.uniform(0.1, 1, size=(N, 2 * N, K, K, K))
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K + 2, 1.0)
    st = tp.DeviceState(vol, b)
    Ha, Hb = b.histogram(ra), b.histogram(rb)
    H2 = b.histogram(ra, range(2, K + 1))
    for _ in range(2):
        for H in (Ha, Hb, H2):
            assert st.energy(H) == tp.energy(vol, b, H)
    for mode in ("fp16x3", "fp16x3_t64", "fp64"):
        b.set_precision(mode)
        assert st.energy(Ha) == tp.energy(vol, b, Ha)
    del Hb
    rc = rng.uniform(0.1, 1, size=(N, 2 * N, K, K, K))
    Hc = b.histogram(rc)
    assert st.energy(Hc) == tp.energy(vol, b, Hc)
    tp.rotate_all_at_once(st, b, Ha, 20, seed=3, ntemp=4)
    moved = tp.SphericalHarmonicsVolume(st.read(), LMAX, K + 2, 1.0)
    assert st.energy(Ha) == tp.energy(moved, b, Ha)


def test_anneal_matches_oracle(tp, oracle):
    """Device-resident annealing vs the oracle driver with the same Philox
    streams: energies within tolerance and identical accept decisions."""
    L, LMAX, N, K = 4, 6, 4, 3
    KMAX = K + 1
    c = random_real_coeff(KMAX, L, LMAX, 12, c00=5.0)
    ob = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(ob.c3_factored(random_real_coeff(K, L, LMAX, 13, c00=5.0)), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    nsteps = 60
    ro = oracle.anneal(ob, ref, c, nsteps, math.pi / 4, 1.01, 0.1, 0.99998, seed=5, ntemp=10)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    st = tp.DeviceState(tp.SphericalHarmonicsVolume(c, LMAX, KMAX, 1.0), b)
    rg = tp.rotate_all_at_once(st, b, ref, nsteps, math.pi / 4, 1.01, 0.1, 0.99998, seed=5, ntemp=10)
    assert rg["E0"] == pytest.approx(ro["E0"], rel=E_TOL)
    assert rg["T0"] == pytest.approx(ro["T0"], rel=1e-8)
    assert np.array_equal(rg["accepted"], ro["accepted"])
    assert rel_err(rg["E_new"], ro["E_new"]) < E_TOL
    assert rel_err(rg["E"], ro["E"]) < E_TOL
    final = st.read()
    assert np.abs(final - ro["coeff"]).max() < 1e-10 * np.abs(ro["coeff"]).max()


def test_intensity_generator_matches_oracle(tp, oracle):
    rng = np.random.default_rng(0)
    atoms = rng.uniform(-3, 3, size=(17, 3))
    KMAX, LMAX, qmax = 4, 10, 2.0
    vol = tp.intensity_from_atoms(atoms, KMAX, qmax, LMAX, 9, 20, 40)
    want = oracle.intensity_from_atoms(atoms, KMAX, qmax / KMAX, LMAX, 9, 20, 40)
    assert np.abs(vol.coeff - want).max() < 1e-11 * np.abs(want).max()


def test_full_size_properties(tp, oracle):
    """At the full shape (L=18, K=26, N=32): columns against the dense
    oracle, and the Gibbs equality E(own normalised c3) = sum tF H(q)."""
    from tp_amd import workload
    cfg = workload.CONFIGS["full"]
    say("full-size: intensity")
    atoms = workload.atoms_in_ball(cfg["natoms"], cfg["radius"], cfg["atoms_seed"])
    vol = tp.intensity_from_atoms(atoms, cfg["KMAX"], cfg["qmax"], cfg["LMAX"], cfg["L"], cfg["ntheta"], cfg["nphi"])
    say("full-size: basis")
    b = tp.HIP_calculate_basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"])
    assert (b.basislen, b.n_lt, b.klength, b.R) == (163153, 505, 3276, 2048)
    c3 = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    say("full-size: oracle columns")
    ob = oracle.Basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"])
    for lo in (0, 1700, 3270):
        assert rel_err(c3[:, lo:lo + 6], ob.c3_dense(vol.coeff, lo, lo + 6)) < C3_TOL
    p = np.maximum(c3, 1e-30)
    ref = c3ref_from_columns(p / p.sum(), cfg["N"], cfg["K"])
    Emin = 0.0
    for i, (k1, k2, k3) in enumerate(ktriples(cfg["K"])):
        q = p[:, i] / p[:, i].sum()
        Emin -= tp.tripletFactor(k1, k2, k3) * float((q * np.log(q)).sum())
    E = tp.energy(vol, b, ref)
    assert abs(E - Emin) / abs(Emin) < E_TOL
    say("full-size: energy ok", E)


def test_basis_from_host_tables(tp, oracle):
    """tp_basis_from_host (CUDA_store_basis) on the oracle's wignerlist/indices
    gives the same c3 as tp_basis_create, and uses the weights it is given."""
    L, LMAX, N, K = 6, 8, 6, 4
    ob = oracle.Basis(L, LMAX, N, K)
    w, ind, _ = ob.tables()
    pac = ob.pacombos()
    c = random_real_coeff(K, L, LMAX, seed=77, c00=3.0)
    b0 = tp.HIP_calculate_basis(L, LMAX, N, K)
    b1 = tp.HIP_store_basis(w, ind, N, L, LMAX, K, pacombos=pac)
    assert (b1.basislen, b1.n_lt, b1.klength) == (b0.basislen, b0.n_lt, b0.klength)
    c0 = tp.FullCorrelation_parallized(c, b0, return_raw=True)
    c1 = tp.FullCorrelation_parallized(c, b1, return_raw=True)
    assert rel_err(c1, c0) < 1e-14
    w2 = w.copy()
    w2[len(w2) // 2] *= 1.5
    b2 = tp.HIP_store_basis(w2, ind, N, L, LMAX, K)
    assert rel_err(tp.FullCorrelation_parallized(c, b2, return_raw=True), c0) > 1e-6
    bad = ind.copy()
    bad[3, 4] += 1   # m1 no longer matches its seanindex
    with pytest.raises(tp.TPError):
        tp.HIP_store_basis(w, bad, N, L, LMAX, K)
