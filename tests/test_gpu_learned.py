"""Device-scored rotation candidates (tp_batch_eval_rotations,
tp_batch_set_state): the scoring step of the learned geodesic proposer
(tp_amd.descent.learned_gradient_runs).  Geodesic candidates exp(-theta G_l)
and the reference's plane move (utilities.jl:180-190), turned and scored on
the device in FP16 hi/lo, against the same rotations built on the host
(eigen-decomposition form) and scored one by one in FP64; the state
hand-over; the driver's bookkeeping (exact energies, c2 invariants, the move
log)."""
import sys

import numpy as np
import pytest

from helpers import c2_invariants, c3ref_from_columns, random_real_coeff

pytestmark = pytest.mark.gpu

MIXED_E_TOL = 1e-6
L, LMAX, N, KB = 8, 10, 8, 6


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


@pytest.fixture(scope="module")
def problem(tp, oracle):
    b = tp.HIP_calculate_basis(L, LMAX, N, KB)
    ob = oracle.Basis(L, LMAX, N, KB)
    truth = random_real_coeff(KB, L, LMAX, 21, c00=5.0)
    p = np.maximum(ob.c3_factored(truth), 1e-30)
    full = c3ref_from_columns(p / p.sum(), N, KB)
    return b, full, truth


def _batch(tp, problem, Ks, nprop):
    b, full, _ = problem
    refs = [np.ascontiguousarray(full[:, :, :K, :K, :K]) for K in Ks]
    starts = [random_real_coeff(KB + 1, L, LMAX, 30 + g, c00=5.0) for g in range(len(Ks))]
    b.set_precision("fp16x3")
    B = tp.AnnealBatch(b, refs, starts, list(range(1, len(Ks) + 1)), nprop=nprop)
    return refs, starts, B


def _fp64(tp, b, c, K, ref):
    cc = np.zeros((KB, c.shape[1]), np.complex128)
    cc[:K] = c[:K]
    prec = b.precision()
    b.set_precision("fp64")
    try:
        return tp.energy(cc, b, ref, range(1, K + 1))
    finally:
        b.set_precision(prec)


def _random_generators(gm, rng, G):
    gens, oms = [], []
    for _ in range(G):
        om = {}
        for l in gm.ls:
            A = rng.normal(size=(2 * l + 1, 2 * l + 1))
            om[l] = 0.5 * (A - A.T)
        oms.append(om)
        gens.append(gm.generators(om)[0])
    return np.stack(gens), oms


def _host_candidate(gm, c, K, om, theta_e, q_e, geo):
    if geo:
        rots = gm.angle_rotations(om, theta_e[None])
    else:
        rots = {}
        for li, l in enumerate(gm.ls):
            n = 2 * l + 1
            from tp_amd.descent import plane_rotation
            R = plane_rotation(q_e[li, :n], q_e[li, n:2 * n], q_e[li, 2 * n])
            rots[l] = (gm.U[l].conj().T @ R.astype(np.complex128) @ gm.U[l])[None]
    return gm.rotate_all(c, rots, K)[0]


def test_eval_rotations_match_fp64_single_evaluations(tp, problem):
    """Two runs (K 5 and 6) x 8 slots, 5 geodesic (one at theta = 0, one at
    theta = 2.9, past the scaling threshold) and 3 plane moves: every energy
    against the FP64 energy of the host-built candidate."""
    from tp_amd import descent
    b = problem[0]
    Ks, nprop, ngeo = (5, 6), 8, 5
    refs, starts, B = _batch(tp, problem, Ks, nprop)
    gm = descent.GeodesicMoves(b)
    rng = np.random.default_rng(3)
    gen, oms = _random_generators(gm, rng, len(Ks))
    nl = len(gm.ls)
    theta = rng.uniform(0.0, 1.5, size=(len(Ks) * nprop, nl))
    theta[0] = 0.0
    theta[nprop + 1] = 2.9
    uvb = np.zeros((len(Ks) * nprop, nl, 2 * (2 * L + 1) + 1))
    for g in range(len(Ks)):
        uvb[g * nprop:(g + 1) * nprop] = descent._plane_draws(nl, gm.ls, [0.4 * (l - 1) for l in gm.ls], rng,
                                                              nprop, L)
    E = B.eval_rotations(ngeo, gen, theta, uvb)
    worst = 0.0
    for g, K in enumerate(Ks):
        for p in range(nprop):
            e = g * nprop + p
            cand = _host_candidate(gm, starts[g], K, oms[g], theta[e], uvb[e], p < ngeo)
            Es = _fp64(tp, b, cand, K, refs[g])
            worst = max(worst, abs(E[e] - Es) / abs(Es))
    say(f"  16 device-turned candidates vs FP64: max rel {worst:.2e}")
    assert worst < MIXED_E_TOL
    E0 = _fp64(tp, b, starts[0], Ks[0], refs[0])
    assert abs(E[0] - E0) / abs(E0) < MIXED_E_TOL


def test_set_state_hands_over_structure_and_energy(tp, problem):
    from tp_amd import descent
    b = problem[0]
    Ks, nprop = (5, 6), 4
    refs, starts, B = _batch(tp, problem, Ks, nprop)
    gm = descent.GeodesicMoves(b)
    rng = np.random.default_rng(8)
    gen, oms = _random_generators(gm, rng, len(Ks))
    nl = len(gm.ls)
    th = np.full(nl, 0.7)
    cand = _host_candidate(gm, starts[1], KB + 1, oms[1], th, None, True)
    Ec = _fp64(tp, b, cand, Ks[1], refs[1])
    B.set_state(1, cand, Ec)
    assert B.group(1)["E"] == Ec
    back = B.coeff(1)
    path = sorted({descent.seanindex(l, m, LMAX) - 1 for l in range(0, L + 1, 2) for m in range(-l, l + 1)})
    assert np.array_equal(back[:Ks[1], path], cand[:Ks[1], path])
    E = B.eval_rotations(nprop, gen, np.zeros((len(Ks) * nprop, nl)), None)
    assert abs(E[nprop] - Ec) / abs(Ec) < MIXED_E_TOL
    assert abs(E[0] - _fp64(tp, b, starts[0], Ks[0], refs[0])) / abs(Ec) < MIXED_E_TOL
    with pytest.raises(tp.TPError):
        bad = np.zeros((len(Ks) * nprop, nl))
        bad[3, 1] = np.nan
        B.eval_rotations(nprop, gen, bad, None)
    with pytest.raises(tp.TPError):
        B.eval_rotations(1, gen, np.zeros((len(Ks) * nprop, nl)), None)   # plane slots need uvb


def test_driver_keeps_exact_energies_and_invariants(tp, problem):
    """learned_gradient_runs without a network (random geodesic angles + plane
    moves), 12 steps of two runs at T = 0: energies never rise, each run's
    energy is the FP64 energy of its coefficients, the rotations keep the c2
    invariants of every shell, and the move log holds unit quaternions."""
    from tp_amd import descent
    b, full, _ = problem
    Ks = (5, 6)
    refs = [np.ascontiguousarray(full[:, :, :K, :K, :K]) for K in Ks]
    starts = [random_real_coeff(KB + 1, L, LMAX, 50 + g, c00=5.0) for g in range(len(Ks))]
    log = descent.MoveLog(50000)
    res, steps = descent.learned_gradient_runs(starts, b, refs, 12, nlearn=12, nrand=4, seed=4, log=log)
    assert steps == 12
    for g, r in enumerate(res):
        Es = [e for _, e in r["history"]]
        assert len(Es) == 13 and all(y <= x for x, y in zip(Es, Es[1:]))
        assert r["E"] < Es[0] and r["accepted"] == sum(r["wins"].values()) > 0
        Es64 = _fp64(tp, b, r["coeff"], Ks[g], refs[g])
        assert abs(r["E"] - Es64) <= 1e-13 * abs(Es64)
        a, c = c2_invariants(starts[g], KB + 1, L, LMAX), c2_invariants(r["coeff"], KB + 1, L, LMAX)
        assert np.abs(a - c).max() < 1e-11 * np.abs(a).max()
        say(f"  run {g} (K={Ks[g]}): E {Es[0]:.6f} -> {r['E']:.6f}, wins {r['wins']}")
    X, T = log.arrays()
    assert X.shape[0] > 0 and np.all(X[:, 190] == -12.0)   # log10 of T = 0, floored
    q = T[:, :4 * len(descent.GeodesicMoves(b).ls)].reshape(T.shape[0], -1, 4)
    assert np.allclose(np.linalg.norm(q, axis=2), 1.0, atol=1e-6)
