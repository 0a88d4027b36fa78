"""The curved Ewald sphere (lambda dq != 0) on the device against the oracle's
dense B.PA form, whose Legendre factors at theta_k = acos(k lambda dq / 4 pi)
are pinned by the exact orientational average in test_oracle.py.  Includes the
reference's own GPU test (tests/test_cuda.jl: the crambin intensity of
test_structure.jl, K = 15, L = 12, N = 16, LMAX = 25, lambda = 2.5, dq of the
intensity), whose bar is rel-L1 < 1e-5; the bar here is the FP64 one, 1e-10."""
import os
import sys

import numpy as np
import pytest

from helpers import c3ref_from_columns, random_real_coeff, rel_err

pytestmark = pytest.mark.gpu
TOL = 1e-10
CRAMBIN = os.path.join(os.path.dirname(__file__), "golden", "crambin_atoms.pdb")


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


CASES = [
    # (L, LMAX, N, K, lambda, dq): the separated-grid contraction with odd m (N <= 32, small L),
    # two k1 halves of the triple products (K = 20), the GEMM contraction (N > 32), the wide
    # separated-grid instance (more than 24 classes per m2: L = 14, and L = 18 with its 40 rows
    # of (m2, cos/sin) at N = 32)
    (4, 6, 4, 3, 2.5, 1.0),
    (6, 9, 5, 5, 1.0, 0.8),
    (4, 6, 32, 4, 2.0, 1.0),
    (6, 8, 4, 20, 0.5, 0.3),
    (4, 6, 40, 3, 2.0, 1.0),
    (14, 16, 24, 6, 1.5, 0.1),
    (18, 19, 32, 5, 2.5, 0.08),
]


@pytest.mark.parametrize("L,LMAX,N,K,lam,dq", CASES)
def test_c3_and_energy_match_oracle(tp, oracle, L, LMAX, N, K, lam, dq):
    say("case", L, LMAX, N, K, lam, dq)
    c = random_real_coeff(K, L, LMAX, seed=L * 10 + K, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq)
    ob = oracle.Basis(L, LMAX, N, K, lam=lam, dq=dq)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    got = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    want = ob.c3_dense(c, nthreads=8)
    err = rel_err(got, want)
    say(f"  c3 {err:.2e}")
    assert err < TOL
    p = np.maximum(want, 1e-30)
    rng = np.random.default_rng(2)
    ref = c3ref_from_columns(rng.poisson(1e4 * p / p.sum() * p.shape[1]).astype(float), N, K)
    ref = np.maximum(ref, 1e-30)
    ref /= ref.sum()
    E, Eo = tp.energy(vol, b, ref), oracle.energy(want, N, ref, K, 1, K)
    say(f"  energy {abs(E - Eo) / abs(Eo):.2e}")
    assert abs(E - Eo) / abs(Eo) < TOL


def test_reference_gpu_test_on_crambin(tp, oracle):
    """tests/test_cuda.jl: calculate_basis(L, LMAX, N, K, 2.5, dq(intensity)) and
    FullCorrelation_parallized of the crambin intensity, GPU against CPU."""
    from tp_amd import structure
    L, LMAX, N, K, lam = 12, 25, 16, 15, 2.5
    _, _, intensity = structure.createSphericalHarmonicsStructure(CRAMBIN, LMAX, 38, 38.0)
    dq = intensity.dq()
    assert dq == pytest.approx(np.pi / 38)
    b = tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq)
    ob = oracle.Basis(L, LMAX, N, K, lam=lam, dq=dq)
    got = tp.FullCorrelation_parallized(intensity, b, return_raw=True)
    want = ob.c3_dense(intensity.coeff[:K], nthreads=8)
    l1 = np.abs(got - want).sum() / np.abs(want).sum()
    say(f"  rel-L1 {l1:.2e}, max {rel_err(got, want):.2e}")
    assert l1 < 1e-5 and rel_err(got, want) < TOL
    # the energy against c3ref = ones(N, 2N, K, K, K), the reference's commented-out check
    ref = np.ones((N, 2 * N, K, K, K))
    E, Eo = tp.energy(intensity, b, ref), oracle.energy(want, N, ref, K, 1, K)
    assert abs(E - Eo) / abs(Eo) < TOL


@pytest.mark.parametrize("L,LMAX,N,K,lam,dq", [(4, 6, 32, 4, 2.0, 1.0), (18, 19, 32, 5, 2.5, 0.08)])
def test_fp16x3_t64_on_the_curved_sphere(tp, oracle, L, LMAX, N, K, lam, dq):
    """FP16X3_T64 at lambda dq != 0: the mixed-precision triple products (FP32 stage 1,
    FP16 hi/lo stage 2) feed the FP64 k_glam and separated-grid contraction.  Bars:
    c3 within 1e-5 of the oracle (the mixed-precision bar) and the energy within 1e-9
    relative; tp_energy_batch (one evaluation after another here) gives the same energies."""
    say("case", L, LMAX, N, K, lam, dq)
    c = random_real_coeff(K, L, LMAX, seed=L * 10 + K, c00=4.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K, lam, dq)
    ob = oracle.Basis(L, LMAX, N, K, lam=lam, dq=dq)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    want = ob.c3_dense(c, nthreads=8)
    p = np.maximum(want, 1e-30)
    rng = np.random.default_rng(3)
    ref = c3ref_from_columns(rng.poisson(1e4 * p / p.sum() * p.shape[1]).astype(float), N, K)
    ref = np.maximum(ref, 1e-30)
    ref /= ref.sum()
    E64 = tp.energy(vol, b, ref)
    b.set_precision("fp16x3_t64")
    got = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    err = rel_err(got, want)
    E = tp.energy(vol, b, ref)
    Eo = oracle.energy(want, N, ref, K, 1, K)
    say(f"  c3 {err:.2e}, energy vs FP64 {abs(E - E64) / abs(E64):.2e}, vs oracle {abs(E - Eo) / abs(Eo):.2e}")
    assert 0 < err < 1e-5
    assert abs(E - Eo) / abs(Eo) < 1e-9
    Eb = tp.energy_batch([vol, vol], [b, b], [ref, ref])
    assert np.array_equal(np.asarray(Eb), np.array([E, E]))


def test_curved_basis_limits(tp):
    b = tp.HIP_calculate_basis(4, 6, 4, 3, 2.5, 1.0)
    for mode in ("fp16x3", "bf16", "triple16"):
        with pytest.raises(RuntimeError, match="lambda"):
            b.set_precision(mode)
    with pytest.raises(RuntimeError, match="acos"):
        tp.HIP_calculate_basis(4, 6, 4, 3, 10.0, 1.0)   # 3 * 10 / (4 pi) > 1
