"""The structure step on the device (tp_density_surface, tp_absolute_square,
tp_structure_from_atoms; createSphericalHarmonicsStructure, structure.jl:124-134)
against the oracle (oracle/structure_oracle.py, pinned in
test_oracle_structure.py), on crambin, the reference's structure-test input.

Bar: density samples within 1e-12 of max |oracle|; coefficient volumes
(density, Fourier transform, intensity) within 1e-10 of max |oracle| per
volume, FP64 throughout."""
import os
import sys

import numpy as np
import pytest

from helpers import random_real_coeff

pytestmark = pytest.mark.gpu
CRAMBIN = os.path.join(os.path.dirname(__file__), "golden", "crambin_atoms.pdb")


@pytest.fixture(scope="module")
def tp():
    import tp_amd
    tp_amd.HIP_init()
    return tp_amd


@pytest.fixture(scope="module")
def so(oracle):
    from oracle import structure_oracle
    return structure_oracle


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _close(got, want, tol):
    err = np.abs(got - want).max() / np.abs(want).max()
    return err <= tol, err


@pytest.mark.parametrize("LMAX,K,rmax", [(6, 5, 10.0), (11, 9, 18.0)])
def test_density_and_structure_match_oracle(tp, so, LMAX, K, rmax):
    from tp_amd import structure
    say(f"case LMAX={LMAX} K={K} rmax={rmax}")
    atoms = structure.atom_table(structure.loadPDB(CRAMBIN))
    ok, err = _close(structure.getDensitySurfaces(atoms, LMAX, K, rmax), so.density_surface(atoms, LMAX, K, rmax), 1e-12)
    say(f"  density samples {err:.2e}")
    assert ok
    d, f, i = structure.createSphericalHarmonicsStructure(CRAMBIN, LMAX, K, rmax)
    wd, wf, wi = so.structure(atoms, LMAX, K, rmax)
    for name, g, w in (("density", d, wd), ("fourier", f, wf), ("intensity", i, wi)):
        ok, err = _close(g.coeff, w, 1e-10)
        say(f"  {name} {err:.2e}")
        assert ok
    assert f.rmax == pytest.approx(np.pi / (rmax / K)) and d.rmax == rmax
    # the step-by-step API gives the same volumes
    d2 = structure.getDensity(atoms, LMAX, K, rmax)
    assert np.allclose(d2.coeff, d.coeff, rtol=0, atol=1e-13 * np.abs(d.coeff).max())


def test_absolute_square_matches_oracle(tp, so):
    from tp_amd import SphericalHarmonicsVolume, structure
    LMAX, K = 12, 4
    c = random_real_coeff(K, LMAX - 1, LMAX, 31, even=False)
    got = structure.absoluteSquare(SphericalHarmonicsVolume(c, LMAX, K, 3.0))
    ok, err = _close(got.coeff, so.absolute_square(c, LMAX), 1e-12)
    say(f"  absoluteSquare {err:.2e}")
    assert ok


def test_crambin_at_the_reference_size(tp, so):
    """LMAX = 35 (the intensity of the histogram-generation runs), K = 38,
    rmax = K as the reference's structure test: every volume against the
    oracle, and Friedel's law (no odd l in the intensity)."""
    from oracle.pyoracle import seanindex
    from tp_amd import structure
    LMAX, K, rmax = 35, 38, 38.0
    atoms = structure.atom_table(structure.loadPDB(CRAMBIN))
    d, f, i = structure.createSphericalHarmonicsStructure(atoms, LMAX, K, rmax)
    wd, wf, wi = so.structure(atoms, LMAX, K, rmax)
    for name, g, w in (("density", d, wd), ("fourier", f, wf), ("intensity", i, wi)):
        ok, err = _close(g.coeff, w, 1e-10)
        say(f"  {name} {err:.2e}")
        assert ok
    odd = [seanindex(l, m, LMAX) - 1 for l in range(1, LMAX, 2) for m in range(-l, l + 1)]
    assert np.abs(i.coeff[:, odd]).max() <= 1e-11 * np.abs(i.coeff).max()


def test_sh_round_trip_above_lmax_32(tp):
    """Synthesis then analysis is the identity on band-limited shells, at the
    LMAX the structure step needs (up to 48)."""
    from tp_amd import SphericalHarmonicsVolume, phases
    for LMAX in (35, 48):
        c = random_real_coeff(3, LMAX - 1, LMAX, LMAX, even=False)
        s = phases.getSurfaceVolume(SphericalHarmonicsVolume(c, LMAX, 3, 1.0))
        back = phases.getSphericalHarmonicsVolume(s, LMAX, 3, 1.0).coeff
        err = np.abs(back - c).max() / np.abs(c).max()
        say(f"  LMAX={LMAX} round trip {err:.2e}")
        assert err <= 1e-11


def test_shell_correlations(tp, so):
    """ISC / FSC (spherical_harmonics.jl:400-423): a structure's own density
    against its intensity and Fourier volume gives 1 on every shell; two
    unrelated volumes give the oracle's surface correlations."""
    from oracle.c2_oracle import surface
    from tp_amd import SphericalHarmonicsVolume, structure
    LMAX, K, rmax = 8, 6, 12.0
    d, f, i = structure.createSphericalHarmonicsStructure(CRAMBIN, LMAX, K, rmax)
    assert np.allclose(structure.ISC(d, i), 1.0, atol=1e-12)
    assert np.allclose(structure.FSC(d, f), 1.0, atol=1e-12)
    a = random_real_coeff(K, LMAX - 1, LMAX, 5, even=False)
    b = random_real_coeff(K, LMAX - 1, LMAX, 6, even=False)
    va, vb = SphericalHarmonicsVolume(a, LMAX, K, 1.0), SphericalHarmonicsVolume(b, LMAX, K, 1.0)
    isc = structure.shell_correlation_ISC(va, vb, range(2, K + 1))
    fsc = structure.shell_correlation_FSC(va, vb)
    for k in range(1, K + 1):
        sa, sb = surface(a[k - 1], LMAX).ravel(), surface(b[k - 1], LMAX).ravel()
        if k >= 2:
            assert isc[k - 2] == pytest.approx(np.corrcoef(sa.real, sb.real)[0, 1], abs=1e-12)
        want = abs(np.vdot(sb, sa)) / (np.linalg.norm(sa) * np.linalg.norm(sb))
        assert fsc[k - 1] == pytest.approx(want, abs=1e-12)


def test_structure_edge_cases(tp):
    """No atoms: zero density, Fourier and intensity volumes; a non-positive
    form factor is an argument error, not a silent NaN."""
    from tp_amd import structure
    d, f, i = structure.createSphericalHarmonicsStructure(np.zeros((0, 5)), 4, 3, 5.0)
    assert not d.coeff.any() and not f.coeff.any() and not i.coeff.any()
    assert not structure.getDensitySurfaces(np.zeros((0, 5)), 4, 3, 5.0).any()
    with pytest.raises(RuntimeError, match="R > 0"):
        structure.getDensitySurfaces(np.array([[0.0, 0.0, 0.0, 0.0, 6.0]]), 4, 3, 5.0)
    # one atom farther than sqrt 5 from every shell point adds nothing
    far = structure.getDensitySurfaces(np.array([[0.0, 0.0, 40.0, 0.7, 6.0]]), 4, 3, 5.0)
    assert not far.any()
