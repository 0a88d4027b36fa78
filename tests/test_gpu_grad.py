"""The energy gradient of the gfx950 path (tp_energy_grad) against the CPU
oracle's chain rule (or_energy_grad, itself pinned by finite differences in
test_oracle.py) and against finite differences of the GPU energy along
rotation generators.

Bar: max |gpu - oracle| <= 1e-9 max |oracle| (FP64 sums in different orders);
the energy returned with the gradient equals tp.energy bit for bit."""
import sys

import numpy as np
import pytest

from helpers import c3ref_from_columns, random_real_coeff, seanindex

pytestmark = pytest.mark.gpu

G_TOL = 1e-9


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(oracle, L, LMAX, N, K, seed):
    c = random_real_coeff(K, L, LMAX, seed, c00=6.0, decay=0.05)
    other = random_real_coeff(K, L, LMAX, seed + 1, c00=6.0, decay=0.05)
    ob = oracle.Basis(L, LMAX, N, K)
    p = np.maximum(ob.c3_factored(other), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    return ob, c, ref


@pytest.mark.parametrize("L,LMAX,N,K,k3hi", [
    (4, 6, 4, 3, None),       # tiny, j range shorter than one 16-block
    (8, 10, 8, 6, None),
    (8, 10, 8, 6, 5),         # K3_range 1:K-1 (a shorter walk)
    (6, 8, 16, 20, None),     # two k1 halves, N = 16
    (18, 19, 32, 5, None),    # l up to 18 and two i blocks (N = 32)
])
def test_energy_grad_matches_oracle(tp, oracle, L, LMAX, N, K, k3hi):
    say(f"case L={L} LMAX={LMAX} N={N} K={K} k3hi={k3hi}")
    ob, c, ref = _case(oracle, L, LMAX, N, K, seed=3 * L + K)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    rng = None if k3hi is None else range(1, k3hi + 1)
    E, g = tp.energy_grad(c, b, ref, rng)
    Eo, go = ob.energy_grad(c, ref, 1, K if k3hi is None else k3hi)
    assert E == tp.energy(c, b, ref, rng)
    assert abs(E - Eo) <= 1e-10 * abs(Eo)
    err = np.abs(g - go).max() / np.abs(go).max()
    say(f"  E {E:.12g}  grad rel err {err:.2e}  |grad| max {np.abs(go).max():.3g}")
    assert np.abs(go).max() > 0
    assert err <= G_TOL


def test_energy_grad_rotation_fd(tp):
    """The derivative along a rotation of one l-block (all shells, real basis)
    from the gradient equals the central difference of the GPU energy."""
    from scipy.linalg import expm
    from tp_amd import workload
    L, LMAX, N, K = 8, 10, 8, 6
    c = random_real_coeff(K, L, LMAX, 5, c00=6.0, decay=0.05)
    other = random_real_coeff(K, L, LMAX, 6, c00=6.0, decay=0.05)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    p = np.maximum(tp.FullCorrelation_parallized(other, b, return_raw=True), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    E, g = tp.energy_grad(c, b, ref)
    rng = np.random.default_rng(2)
    for l in (2, 6, 8):
        n = 2 * l + 1
        A = rng.normal(size=(n, n))
        A = A - A.T
        eps = 1e-5

        def rot(t):
            return workload.rotate_l(c, l, expm(t * A), LMAX)

        fd = (tp.energy(rot(eps), b, ref) - tp.energy(rot(-eps), b, ref)) / (2 * eps)
        d = (rot(eps) - rot(-eps)) / (2 * eps)
        an = float(np.real(np.vdot(g, d)))
        say(f"  l={l}: analytic {an:.10g}  fd {fd:.10g}")
        assert an == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(E))


@pytest.mark.parametrize("L,LMAX,N,K,lam,dq", [
    (8, 10, 8, 6, 2.5, 0.3),      # separated grid with odd m
    (14, 16, 16, 6, 1.5, 0.3),    # the wide layout (> 24 classes per m2), two j blocks
    (18, 19, 32, 5, 2.5, 0.08),   # the wide layout at L = 18, N = 32
])
def test_energy_grad_curved_sphere_fd(tp, L, LMAX, N, K, lam, dq):
    """The gradient on the curved Ewald sphere (lambda dq != 0: the contraction's
    backward in the odd-m layout, k_glam run backwards) against central
    differences of the GPU energy (pinned to the oracle's dense form at
    lambda dq != 0 in test_gpu_ewald.py; the oracle's chain rule covers lambda
    = 0 only): along rotations of l-blocks and along a random direction with
    the real-intensity structure.  The energy returned equals tp.energy."""
    from scipy.linalg import expm
    from tp_amd import workload
    say(f"case L={L} N={N} K={K} lambda={lam} dq={dq}")
    c = random_real_coeff(K, L, LMAX, 5, c00=6.0, decay=0.05)
    other = random_real_coeff(K, L, LMAX, 6, c00=6.0, decay=0.05)
    b = tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq)
    p = np.maximum(tp.FullCorrelation_parallized(other, b, return_raw=True), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    E, g = tp.energy_grad(c, b, ref)
    assert E == tp.energy(c, b, ref)
    assert np.abs(g).max() > 0
    rng = np.random.default_rng(2)
    eps = 1e-5
    for l in (2, L):
        n = 2 * l + 1
        A = rng.normal(size=(n, n))
        A = A - A.T

        def rot(t):
            return workload.rotate_l(c, l, expm(t * A), LMAX)

        fd = (tp.energy(rot(eps), b, ref) - tp.energy(rot(-eps), b, ref)) / (2 * eps)
        d = (rot(eps) - rot(-eps)) / (2 * eps)
        an = float(np.real(np.vdot(g, d)))
        say(f"  rotation l={l}: analytic {an:.10g}  fd {fd:.10g}")
        assert an == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(E))
    d = random_real_coeff(K, L, LMAX, 11, c00=1.0, decay=0.05)
    fd = (tp.energy(c + eps * d, b, ref) - tp.energy(c - eps * d, b, ref)) / (2 * eps)
    an = float(np.real(np.vdot(g, d)))
    say(f"  random direction: analytic {an:.10g}  fd {fd:.10g}")
    assert an == pytest.approx(fd, rel=1e-5, abs=1e-7 * abs(E))


def test_gradient_moves_beat_random_walk(tp):
    """Gradient-guided moves (tp_amd.descent) only ever lower the energy at
    T = 0 and, for the same number of energy evaluations (a gradient counted
    as three), end far lower than the reference's random-walk annealing."""
    from tp_amd import descent, workload
    L, LMAX, N, K = 8, 10, 8, 6
    truth = random_real_coeff(K, L, LMAX, 31, c00=6.0, decay=0.05)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    p = np.maximum(tp.FullCorrelation_parallized(truth, b, return_raw=True), 1e-30)
    ref = c3ref_from_columns(p / p.sum(), N, K)
    g = np.random.default_rng(7)
    start = truth.copy()
    for l in range(2, L + 1, 2):
        start = workload.rotate_l(start, l, workload.haar_rotation(2 * l + 1, g), LMAX)
    E0 = tp.energy(start, b, ref)
    nsteps, ncand = 12, 8
    r = descent.gradient_anneal(start, b, ref, nsteps, ncand=ncand)
    Es = [e for _, e in r["history"]]
    say(f"  gradient: {Es[0]:.6f} -> {Es[-1]:.6f} in {len(Es) - 1} steps, step lengths "
        f"{np.round(r['step_lengths'], 3)}")
    assert all(Es[i + 1] <= Es[i] for i in range(len(Es) - 1))
    assert Es[0] == E0
    evals = nsteps * (ncand + 3)
    st = tp.DeviceState(tp.SphericalHarmonicsVolume(start, LMAX, K, float(K)), b)
    plain = tp.rotate_all_at_once(st, b, ref, evals, np.pi / 4, 1.01, 0.1, 0.99998, seed=3, ntemp=10)
    say(f"  random walk, {evals} steps: {plain['E0']:.6f} -> {plain['E_final']:.6f}")
    assert E0 - Es[-1] > 2.0 * (E0 - plain["E_final"])
