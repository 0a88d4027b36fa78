"""Host side of the structure step (tp_amd.structure), no GPU: the PDB
reader at the reference's columns, centring, and the form-factor table."""
import os

import numpy as np
import pytest

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
CRAMBIN = os.path.join(GOLDEN, "crambin_atoms.pdb")


def test_load_pdb_reads_the_reference_columns():
    from tp_amd.structure import loadPDB
    atoms = loadPDB(CRAMBIN, center=False)
    assert len(atoms) == 670
    a = atoms[0]   # "ATOM      1  N   THR A   1      -7.674  -4.415  -3.515  1.00  9.99           N"
    assert a.index == 1 and a.special_type == " N  " and a.t == "N" and a.group == "THR"
    assert np.array_equal(a.pos, [-7.674, -4.415, -3.515]) and a.bfactor == 9.99
    counts = {}
    for x in atoms:
        counts[x.t] = counts.get(x.t, 0) + 1
    assert counts == {"C": 211, "H": 329, "N": 57, "O": 67, "S": 6}


def test_center_structure_is_the_unweighted_mean():
    from tp_amd.structure import loadPDB
    raw = np.array([a.pos for a in loadPDB(CRAMBIN, center=False)])
    cen = np.array([a.pos for a in loadPDB(CRAMBIN)])
    assert np.allclose(cen, raw - raw.mean(0), atol=1e-12)
    assert np.abs(cen.mean(0)).max() < 1e-12


def test_atom_table_form_factors_and_type_one():
    from tp_amd.structure import Atom, atom_table, atomfactorlist
    mk = lambda t: Atom(1, " X  ", np.array([1.0, 2.0, 3.0]), t, " X  ", 0.0, "GLY")
    tab = atom_table([mk("C"), mk("1"), mk("S"), mk("H")])
    assert tab.shape == (3, 5)   # type "1" carries no density (structure.jl:78)
    assert np.array_equal(tab[:, 3:], [atomfactorlist["C"], atomfactorlist["S"], atomfactorlist["H"]])
    with pytest.raises(KeyError):
        atom_table([mk("X")])


def test_correlation_rules():
    """cor_FSC: |sum a conj b| / (|a| |b|), 0 at zero norm, invariant under a
    complex phase; cor_ISC: Pearson of the real parts (utilities.jl:45-59)."""
    from tp_amd.structure import cor_FSC, cor_ISC
    rng = np.random.default_rng(1)
    a = rng.normal(size=50) + 1j * rng.normal(size=50)
    assert cor_FSC(a, np.zeros(50, complex)) == 0.0
    assert cor_FSC(a, 3.0 * np.exp(0.7j) * a) == pytest.approx(1.0, abs=1e-14)
    assert cor_ISC(a, 2.0 * a + 5.0) == pytest.approx(1.0, abs=1e-14)
    assert cor_ISC(a, -a) == pytest.approx(-1.0, abs=1e-14)
