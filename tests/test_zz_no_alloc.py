"""tp_energy_noalloc (every per-evaluation buffer in the library's static
scratch array, no allocation) against tp_energy, through the C ABI: the same
kernels in the same order, so the energies must be identical bit for bit.
Two calls in a row, the second on a smaller basis (the array keeps the first
call's values), in FP64 and in the mixed-precision modes, and a basis too
large for the array refused with TP_E_UNSUPPORTED."""
import ctypes as C
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


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _case(tp, L, LMAX, N, K, seed, ld_extra=0):
    b = tp.HIP_calculate_basis(L, LMAX, N, K)
    c = random_real_coeff(K, L, LMAX, seed, c00=4.0)
    if ld_extra:   # a caller's vectors longer than LMAX^2 (ld > LMAX^2)
        c = np.concatenate([c, np.full((K, ld_extra), 7.0 + 0j)], axis=1)
    ref = np.random.default_rng(seed + 1).uniform(0.2, 1.0, size=(N, 2 * N, K, K, K))
    return b, np.ascontiguousarray(c), ref


def _both(tp, b, c, ref):
    L = tp.lib()
    h = b.ref(ref)
    e1, e2 = C.c_double(), C.c_double()
    tp.check(L.tp_energy(b.ctx.h, b.h, h, c.ctypes.data, c.shape[1], 0, C.byref(e1)), "tp_energy")
    tp.check(L.tp_energy_noalloc(b.ctx.h, b.h, h, c.ctypes.data, c.shape[1], 0, C.byref(e2)), "tp_energy_noalloc")
    return e1.value, e2.value


@pytest.mark.parametrize("prec", ["fp64", "fp16x3", "triple16"])
def test_noalloc_equals_tp_energy_two_calls(tp, prec):
    cases = [(8, 25, 16, 8, 3, 0), (4, 6, 4, 3, 5, 11)]   # the second call smaller, with ld > LMAX^2
    for (L_, LMAX, N, K, seed, extra) in cases:
        say("noalloc", prec, L_, LMAX, N, K)
        b, c, ref = _case(tp, L_, LMAX, N, K, seed, extra)
        b.set_precision(prec)
        e1, e2 = _both(tp, b, c, ref)
        say(f"  tp_energy {e1!r} noalloc {e2!r}")
        assert np.isfinite(e1) and e1 == e2
        # and again: the array now holds this call's values
        e3 = _both(tp, b, c, ref)[1]
        assert e3 == e1


def test_noalloc_full_size_and_ragged(tp):
    """The full-size shape (L=18, K=26, N=32: about 20 MiB of scratch) then a
    ragged one (R and klength not tile multiples)."""
    for (L_, LMAX, N, K, seed) in [(18, 25, 32, 26, 9), (6, 9, 5, 5, 10)]:
        say("noalloc size", L_, LMAX, N, K)
        b, c, ref = _case(tp, L_, LMAX, N, K, seed)
        e1, e2 = _both(tp, b, c, ref)
        assert e1 == e2


def test_noalloc_refuses_a_basis_beyond_its_array(tp):
    b, c, ref = _case(tp, 18, 25, 4, 40, 4)   # As alone: 512 x 11,520 doubles = 47 MB
    L = tp.lib()
    h = b.ref(ref)
    e = C.c_double()
    rc = L.tp_energy_noalloc(b.ctx.h, b.h, h, c.ctypes.data, c.shape[1], 0, C.byref(e))
    assert rc == -5 and b"scratch" in L.tp_last_error()
    # tp_energy still serves it
    e1 = C.c_double()
    tp.check(L.tp_energy(b.ctx.h, b.h, h, c.ctypes.data, c.shape[1], 0, C.byref(e1)), "tp_energy")
    assert np.isfinite(e1.value)
