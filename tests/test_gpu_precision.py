"""Mixed-precision contraction (TP_PREC_FP16X3 / TP_PREC_FP16X3_T64 /
TP_PREC_BF16 / TP_PREC_TRIPLE16) against the FP64 path, which itself matches the oracle to ~1e-15 (test_gpu_parity.py).

Bar: c3 relative error <= 1e-5 against FP64 for the compensated
FP16 form.  c3 error per k-triple slice = max_r |c3 - c3_64| / max_r |c3_64|."""
import numpy as np
import pytest

from helpers import c3ref_from_columns, random_real_coeff

pytestmark = pytest.mark.gpu

C3_TOL_FP16X3 = 1e-5


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


def slice_err(c, ref):
    return float((np.abs(c - ref).max(axis=0) / np.abs(ref).max(axis=0)).max())


@pytest.mark.parametrize("L,LMAX,N,K", [(4, 6, 4, 3), (8, 25, 16, 8), (6, 8, 5, 20)])
def test_fp16x3_c3_and_energy(tp, L, LMAX, N, K):
    c = random_real_coeff(K, L, LMAX, seed=31 + K, c00=4.0, decay=0.1)
    vol = tp.SphericalHarmonicsVolume(c, LMAX, K, 1.0)
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    c64 = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    p = np.maximum(c64, 1e-30)
    ref = c3ref_from_columns(np.random.default_rng(2).poisson(1e6 * p / p.sum()).astype(float) + 1e-30, N, K)
    E64 = tp.energy(vol, b, ref)
    b.set_precision("fp16x3")
    c16 = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    E16 = tp.energy(vol, b, ref)
    b.set_precision("bf16")
    cbf = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    b.set_precision("triple16")
    ct16 = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    Et16 = tp.energy(vol, b, ref)
    b.set_precision("fp16x3_t64")
    c64t = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    E64t = tp.energy(vol, b, ref)
    b.set_precision("fp64")
    again = tp.FullCorrelation_parallized(vol, b, return_raw=True)
    assert np.array_equal(again, c64)            # switching modes leaves the FP64 path untouched
    e16, ebf = slice_err(c16, c64), slice_err(cbf, c64)
    print(f"L={L} K={K}: fp16x3 slice err {e16:.2e}, bf16 {ebf:.2e}, |dE| {abs(E16 - E64):.2e} of {E64:.6g}")
    assert e16 < C3_TOL_FP16X3
    assert ebf > 10 * e16                         # one BF16 product is not enough for 1e-5
    assert ebf < 5e-2
    assert abs(E16 - E64) / abs(E64) < 1e-6
    et = slice_err(ct16, c64)                    # triple products alone in compensated FP16
    print(f"   triple16 slice err {et:.2e}, |dE| {abs(Et16 - E64):.2e}")
    assert et < C3_TOL_FP16X3
    assert abs(Et16 - E64) / abs(E64) < 1e-6
    e64t = slice_err(c64t, c64)                  # the contraction alone in FP16 hi/lo
    print(f"   fp16x3_t64 slice err {e64t:.2e}, |dE| {abs(E64t - E64):.2e}")
    assert e64t < C3_TOL_FP16X3
    assert abs(E64t - E64) / abs(E64) < 1e-6


def test_fp16x3_full_size(tp):
    from tp_amd import workload
    cfg = workload.CONFIGS["full"]
    atoms = workload.atoms_in_ball(cfg["natoms"], cfg["radius"], cfg["atoms_seed"])
    vol = tp.intensity_from_atoms(atoms, cfg["KMAX"], cfg["qmax"], cfg["LMAX"], cfg["L"], cfg["ntheta"], cfg["nphi"])
    b = tp.HIP_calculate_basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"])
    errs = []
    for seed in range(3):
        st = workload.random_start_structure(vol, cfg["L"], seed)
        b.set_precision("fp64")
        c64 = tp.FullCorrelation_parallized(st, b, return_raw=True)
        b.set_precision("fp16x3")
        c16 = tp.FullCorrelation_parallized(st, b, return_raw=True)
        b.set_precision("fp16x3_t64")
        c16t = tp.FullCorrelation_parallized(st, b, return_raw=True)
        errs.append((slice_err(c16, c64), slice_err(c16t, c64)))
    b.set_precision("fp64")
    print("full-size fp16x3 / fp16x3_t64 slice errors:", errs)
    assert max(max(e) for e in errs) < C3_TOL_FP16X3
