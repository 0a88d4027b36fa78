"""Batched evaluation path (tp_energy_batch, tp_batch_*) and the FP64 Wigner-D
rotation (tp_rotate) on the GPU.

Pins:
* rotation: f'(r) = f(R^-1 r) evaluated through the oracle's sphPlm (the
  reference's spherical harmonics) at random points -- this fixes the real
  Wigner-D matrices completely; one quaternion for all l leaves c3 unchanged
  (SO(3) invariance, checked on the FP64 path against the unrotated c3);
  per-l rotations keep every two-photon sum_m conj(c_k1) c_k2.
* batch energies: equal to single evaluations of the same coefficients in the
  same precision mode (same kernels per column; only the order of the final
  sums differs), tolerance 1e-10 relative; and to FP64 within the
  mixed-precision bar.
"""
import numpy as np
import pytest

from helpers import (c2_invariants, c3ref_from_columns, quat_to_rot, random_quats, random_real_coeff, rel_err,
                     sh_eval)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def test_rotate_equivariance(tp, oracle):
    L, LMAX, K = 8, 10, 3
    c = random_real_coeff(K, L, LMAX, seed=5)
    b = tp.HIP_calculate_basis(L, LMAX, 4, K)
    rng = np.random.default_rng(11)
    q = random_quats(rng, L // 2)
    out = tp.rotate(c, b, q)
    for li, l in enumerate(range(2, L + 1, 2)):
        Rm = quat_to_rot(q[li])
        for k in range(K):
            for _ in range(4):
                v = rng.normal(size=3)
                v /= np.linalg.norm(v)
                lhs = sh_eval(oracle, out[k], l, LMAX, v)
                rhs = sh_eval(oracle, c[k], l, LMAX, Rm.T @ v)
                assert abs(lhs - rhs) < 1e-12 * (1 + abs(rhs)), (l, k, lhs, rhs)
    # l = 0 untouched
    assert np.array_equal(out[:, 0], c[:, 0])


def test_rotate_invariances(tp):
    L, LMAX, N, K = 8, 25, 8, 5
    c = random_real_coeff(K, L, LMAX, seed=8, c00=4.0, decay=0.1)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    c3 = tp.FullCorrelation_parallized(c, b, return_raw=True)
    rng = np.random.default_rng(3)
    qall = random_quats(rng, 1)[0]
    rot_all = tp.rotate(c, b, qall)
    assert rel_err(tp.FullCorrelation_parallized(rot_all, b, return_raw=True), c3) < 1e-12
    qper = random_quats(rng, L // 2)
    rot_l = tp.rotate(c, b, qper)
    c2a, c2b = c2_invariants(c, K, L, LMAX), c2_invariants(rot_l, K, L, LMAX)
    assert np.abs(c2a - c2b).max() < 1e-12 * np.abs(c2a).max()
    # per-l rotations change c3 (they are real moves)
    assert rel_err(tp.FullCorrelation_parallized(rot_l, b, return_raw=True), c3) > 1e-4


def _hist(tp, c, L, LMAX, N, K, seed, lam=0.0, dq=0.0):
    b = tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq)
    c64 = tp.FullCorrelation_parallized(c, b, return_raw=True)
    p = np.maximum(c64, 1e-30)
    cnt = np.random.default_rng(seed).poisson(1e6 * p / p.sum()).astype(float) + 1e-30
    return c3ref_from_columns(cnt, N, K)


@pytest.mark.parametrize("prec,N,L,lam", [("fp16x3", 8, 8, 0.0), ("fp64", 8, 8, 0.0), ("fp16x3_t64", 8, 8, 0.0),
                                         ("fp16x3_t64", 7, 8, 0.0),    # odd N: the full (alpha, beta) grid
                                         ("fp16x3_t64", 8, 8, 2.5),    # curved Ewald sphere (k_glam)
                                         ("fp16x3_t64", 8, 14, 1.5)])  # ... with the wide separated grid
def test_energy_batch_matches_single(tp, prec, N, L, lam):
    LMAX = 25
    dq = 0.3 if lam else 0.0   # |K lambda dq| <= 4 pi at K = 12
    Ks = [3, 9, 5, 12]
    starts = [random_real_coeff(K, L, LMAX, seed=40 + K, c00=4.0, decay=0.1) for K in Ks]
    refs = [_hist(tp, starts[i], L, LMAX, N, K, 70 + i, lam, dq) for i, K in enumerate(Ks)]
    bases = [tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq) for K in Ks]
    for b in bases:
        b.set_precision(prec)
    rng = np.random.default_rng(1)
    for rep in range(3):   # the second and third calls replay the cached graph
        cs = [tp.rotate(starts[i], bases[i], random_quats(rng, L // 2)) for i in range(len(Ks))]
        Eb = tp.energy_batch(cs, bases, refs)
        Es = np.array([tp.energy(cs[i], bases[i], refs[i]) for i in range(len(Ks))])
        print(prec, rep, Eb, np.abs(Eb - Es) / np.abs(Es))
        if prec == "fp64":
            assert np.array_equal(Eb, Es)
        else:
            assert np.all(np.abs(Eb - Es) <= 1e-10 * np.abs(Es))
    for b in bases:
        b.set_precision("fp64")
    E64 = np.array([tp.energy(cs[i], bases[i], refs[i]) for i in range(len(Ks))])
    assert np.all(np.abs(Eb - E64) <= 1e-6 * np.abs(E64))


def _anneal_setup(tp, Ks, nprop, hist_cap=0, L=8, LMAX=25, N=8, prec="fp16x3", lam=0.0, dq=0.0):
    master = tp.HIP_calculate_basis(L, LMAX, N, max(Ks), lam, dq)
    master.set_precision(prec)
    bases = [tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq) for K in Ks]
    truth = [random_real_coeff(K, L, LMAX, seed=90 + K, c00=4.0, decay=0.1) for K in Ks]
    refs = [_hist(tp, truth[i], L, LMAX, N, K, 20 + i, lam, dq) for i, K in enumerate(Ks)]
    rng = np.random.default_rng(4)
    starts = [tp.rotate(truth[i], bases[i], random_quats(rng, L // 2)) for i in range(len(Ks))]
    B = tp.AnnealBatch(master, refs, starts, seeds=[1000 + g for g in range(len(Ks))], nprop=nprop,
                       hist_cap=hist_cap)
    return master, bases, refs, starts, B


@pytest.mark.parametrize("prec,lam", [("fp16x3", 0.0), ("fp16x3_t64", 0.0), ("fp16x3_t64", 2.5)])
def test_batch_eval_quats_matches_single(tp, prec, lam):
    """A batch's energies equal single evaluations in the same mode (the
    batch keeps its mode when the basis changes afterwards); lam != 0: the
    curved Ewald sphere."""
    Ks, nprop = [4, 9, 6], 4
    L = 8
    master, bases, refs, starts, B = _anneal_setup(tp, Ks, nprop, prec=prec, lam=lam, dq=0.3 if lam else 0.0)
    master.set_precision("fp64")
    rng = np.random.default_rng(9)
    q = np.stack([random_quats(rng, L // 2) for _ in range(len(Ks) * nprop)])
    E = B.eval_quats(q)
    for b in bases:
        b.set_precision(prec)
    for g, K in enumerate(Ks):
        for p in range(nprop):
            e = g * nprop + p
            c = tp.rotate(starts[g], bases[g], q[e])
            Es = tp.energy(c, bases[g], refs[g])
            assert abs(E[e] - Es) <= 1e-10 * abs(Es), (g, p, E[e], Es)


def test_batch_sweeps_curved_sphere(tp):
    """Annealing sweeps of a batch on the curved Ewald sphere (lambda dq != 0,
    FP16X3_T64: k_glam over the batch's columns): every run's energy equals a
    single evaluation of its state, and the runs descend."""
    Ks, nprop, nsw = [5, 9, 7], 8, 10
    master, bases, refs, starts, B = _anneal_setup(tp, Ks, nprop, hist_cap=nsw, prec="fp16x3_t64", lam=2.5,
                                                   dq=0.3)
    B.probe(0.1)
    g0 = [B.group(g) for g in range(len(Ks))]
    for b in bases:
        b.set_precision("fp16x3_t64")
    for g in range(len(Ks)):
        Es = tp.energy(starts[g], bases[g], refs[g])
        assert abs(g0[g]["E"] - Es) <= 1e-10 * abs(Es)
    B.run(nsw, decay=0.999, factor=1.01)
    for g in range(len(Ks)):
        st = B.group(g)
        assert st["nsweep"] == nsw
        Es = tp.energy(B.coeff(g), bases[g], refs[g])
        assert abs(st["E"] - Es) <= 1e-10 * abs(Es)
        assert st["E"] <= g0[g]["E"] + 5 * g0[g]["T"]


@pytest.mark.parametrize("family", ["plane", "quat"])
def test_batch_sweeps(tp, family):
    Ks, nprop, nsw = [5, 9, 7], 16, 25
    L, LMAX = 8, 25
    master, bases, refs, starts, B = _anneal_setup(tp, Ks, nprop, hist_cap=nsw)
    B.set_family(family)
    cfg = B.autotune(reps=2)
    assert cfg["tile"] in [(128, 64), (64, 64), (128, 128), (64, 128)] and cfg["order"] in (0, 1)
    assert cfg["schedule"] in ("rows", "walk", "xcdrows")
    B.probe(0.1)
    g0 = [B.group(g) for g in range(len(Ks))]
    for b in bases:
        b.set_precision("fp16x3")
    for g in range(len(Ks)):
        Es = tp.energy(starts[g], bases[g], refs[g])
        assert abs(g0[g]["E"] - Es) <= 1e-10 * abs(Es)
        assert g0[g]["T"] > 0
    B.run(nsw, decay=0.999, factor=1.01)
    hist = B.history()
    for g in range(len(Ks)):
        st = B.group(g)
        assert st["nsweep"] == nsw
        assert st["T"] == pytest.approx(g0[g]["T"] * 0.999 ** nsw, rel=1e-12)
        assert hist[-1, g] == st["E"]
        c = B.coeff(g)
        Es = tp.energy(c, bases[g], refs[g])
        assert abs(st["E"] - Es) <= 1e-10 * abs(Es)
        # moves are rotations of l-blocks: the two-photon sums are kept
        a, b2 = c2_invariants(starts[g], Ks[g], L, LMAX), c2_invariants(c, Ks[g], L, LMAX)
        assert np.abs(a - b2).max() < 1e-11 * np.abs(a).max()
        assert st["E"] <= g0[g]["E"] + 5 * g0[g]["T"]
    # deterministic: a second batch with the same seeds follows the same path
    _, _, _, _, B2 = _anneal_setup(tp, Ks, nprop, hist_cap=nsw)
    B2.set_family(family)
    B2.probe(0.1)
    B2.run(nsw, decay=0.999, factor=1.01)
    # (its contraction tile may differ from the autotuned one: same values up to summation order)
    np.testing.assert_allclose(B2.history(), hist, rtol=1e-10)


@pytest.mark.parametrize("prec,ncand", [("fp16x3", 12), ("fp16x3_t64", 4)])
def test_batch_every_contraction_candidate(tp, prec, ncand):
    """Every contraction candidate of the autotuner, forced in turn, gives the
    single evaluations' energies (N = 32: 2048 rows, so the XCD-partitioned
    schedule applies to every tile of the dense form; the factored form's
    candidates are k_hgemm_g's column and row-group shapes)."""
    Ks, nprop, L, N = [4, 7, 5], 4, 4, 32
    master, bases, refs, starts, B = _anneal_setup(tp, Ks, nprop, L=L, LMAX=6, N=N, prec=prec)
    rng = np.random.default_rng(12)
    q = np.stack([random_quats(rng, L // 2) for _ in range(len(Ks) * nprop)])
    for b in bases:
        b.set_precision(prec)
    Es = []
    for g in range(len(Ks)):
        for p in range(nprop):
            Es.append(tp.energy(tp.rotate(starts[g], bases[g], q[g * nprop + p]), bases[g], refs[g]))
    Es = np.array(Es)
    for cand in range(ncand):
        print(prec, "candidate", cand, flush=True)
        B.set_contraction(cand)
        E = B.eval_quats(q)
        assert np.all(np.abs(E - Es) <= 1e-10 * np.abs(Es)), (cand, np.abs(E - Es).max())
