"""Structures: atoms from PDB files and their spherical-harmonics density,
Fourier transform and intensity (structure.jl:1-134) on the device.

  loadPDB(filename, center=True)              structure.jl:30-54
  centerStructure(atomlist)                   structure.jl:57-71
  getDensity(atomlist, LMAX, KMAX, rmax)      spherical_harmonics.jl:234-244 (tp_density_surface + tp_sh_coeff)
  absoluteSquare(volume)                      spherical_harmonics.jl:393-397 (tp_absolute_square)
  createSphericalHarmonicsStructure(pdb_or_atoms, LMAX, KMAX, rmax)
                                              structure.jl:124-134 (tp_structure_from_atoms)
  shell_correlation / _ISC / _FSC, ISC, FSC    spherical_harmonics.jl:400-423: per-shell
                                              correlations of surfaces synthesised on the device

The atoms' form factors (R, B) come from atomfactorlist (structure.jl:8-17).
As in getDensity_cartesian, atoms of type "1" add no density; any other type
missing from the table raises KeyError, as the reference's dictionary lookup
does.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import HIP_init, SphericalHarmonicsVolume, check, lib

# element -> (R, B): Gaussian width and weight of the atom's electron cloud
atomfactorlist = {
    "C": (0.7, 6.0),
    "N": (0.65, 7.0),
    "S": (1.0, 16.0),
    "O": (0.6, 8.0),
    "default": (0.7, 6.0),
    "P": (0.65, 15.0),
    "H": (0.3, 1.0),
    "D": (0.3, 1.0),
}


@dataclass
class Atom:
    """atomType (structure.jl:20-28)."""
    index: int
    special_type: str
    pos: np.ndarray
    t: str
    specific_type: str
    bfactor: float
    group: str


def loadPDB(filename, center=True):
    """ATOM records of a PDB file, at the reference's fixed columns: serial
    digits 9-11, name 13-16, x/y/z 31-38/39-46/47-54, B-factor 62-66,
    element 78 (1-based, inclusive)."""
    atoms = []
    with open(filename) as f:
        for line in f:
            line = line.rstrip("\n")
            if len(line) > 4 and line[:4] == "ATOM":
                atoms.append(Atom(int(line[8:11]), line[12:16],
                                  np.array([float(line[30:38]), float(line[38:46]), float(line[46:54])]),
                                  line[77:78], line[12:16], float(line[61:66]), line[17:20]))
    return centerStructure(atoms) if center else atoms


def centerStructure(atomlist):
    """Shift every atom by the mean position (unweighted, all atoms)."""
    if not atomlist:
        return atomlist
    center = np.zeros(3)
    for a in atomlist:
        center = center + a.pos
    center = center / len(atomlist)
    for a in atomlist:
        a.pos = a.pos - center
    return atomlist


def atom_table(atomlist):
    """(n, 5) rows (x, y, z, R, B) of the atoms that carry density."""
    rows = []
    for a in atomlist:
        if a.t == "1":
            continue
        R, B = atomfactorlist[a.t]
        rows.append([a.pos[0], a.pos[1], a.pos[2], R, B])
    return np.ascontiguousarray(np.array(rows, dtype=np.float64).reshape(-1, 5))


def _atoms(x):
    if isinstance(x, str):
        x = loadPDB(x)
    if isinstance(x, np.ndarray):
        a = np.ascontiguousarray(x, dtype=np.float64)
        if a.ndim != 2 or a.shape[1] != 5:
            raise ValueError("atom rows must be (n, 5): x, y, z, R, B")
        return a
    return atom_table(x)


def getDensitySurfaces(atomlist, LMAX, KMAX, rmax, ctx=None):
    """(KMAX, (2 LMAX)^2) density samples of getDensity's shells (before the transform)."""
    ctx = ctx or HIP_init()
    a = _atoms(atomlist)
    out = np.zeros((KMAX, 4 * LMAX * LMAX), np.complex128)
    check(lib().tp_density_surface(ctx.h, a.ctypes.data if len(a) else None, len(a), LMAX, KMAX, float(rmax),
                                   out.ctypes.data), "tp_density_surface")
    return out


def getDensity(atomlist, LMAX, KMAX, rmax, ctx=None):
    from .phases import getSphericalHarmonicsVolume
    return getSphericalHarmonicsVolume(getDensitySurfaces(atomlist, LMAX, KMAX, rmax, ctx), LMAX, KMAX, rmax, ctx)


def absoluteSquare(volume, ctx=None):
    ctx = ctx or HIP_init()
    c = np.ascontiguousarray(volume.coeff, dtype=np.complex128)
    out = np.zeros_like(c)
    check(lib().tp_absolute_square(ctx.h, volume.LMAX, volume.KMAX, c.ctypes.data, c.shape[1], out.ctypes.data),
          "tp_absolute_square")
    return SphericalHarmonicsVolume(out, volume.LMAX, volume.KMAX, volume.rmax)


def createSphericalHarmonicsStructure(pdb_or_atoms, LMAX, KMAX, rmax, ctx=None):
    """(density, fourier, intensity) SphericalHarmonicsVolumes; fourier and
    intensity carry rmax = qmax = pi / (rmax / KMAX), as forward sets it."""
    ctx = ctx or HIP_init()
    a = _atoms(pdb_or_atoms)
    n = LMAX * LMAX
    d, f, i = (np.zeros((KMAX, n), np.complex128) for _ in range(3))
    check(lib().tp_structure_from_atoms(ctx.h, a.ctypes.data if len(a) else None, len(a), LMAX, KMAX, float(rmax),
                                        d.ctypes.data, f.ctypes.data, i.ctypes.data, n), "tp_structure_from_atoms")
    q = np.pi / (rmax / KMAX)
    return (SphericalHarmonicsVolume(d, LMAX, KMAX, rmax), SphericalHarmonicsVolume(f, LMAX, KMAX, q),
            SphericalHarmonicsVolume(i, LMAX, KMAX, q))


# ---- shell correlations (spherical_harmonics.jl:400-423, utilities.jl:45-59)

def cor_ISC(a, b):
    """Pearson correlation of the real parts (utilities.jl:45-47)."""
    return float(np.corrcoef(np.real(a), np.real(b))[0, 1])


def cor_FSC(a, b):
    """|sum a conj(b)| / (|a| |b|), 0 when either norm is at most eps (utilities.jl:50-59)."""
    na, nb = np.linalg.norm(a), np.linalg.norm(b)
    if na > np.finfo(float).eps and nb > np.finfo(float).eps:
        return float(abs(np.sum(a * np.conj(b)) / (na * nb)))
    return 0.0


def shell_correlation(volume1, volume2, K_range=None, cor_method=cor_FSC, ctx=None):
    """Per-shell correlation of the two volumes' surfaces (synthesised on the device)."""
    from .phases import getSurfaceVolume
    s1, s2 = getSurfaceVolume(volume1, ctx), getSurfaceVolume(volume2, ctx)
    if s1.shape != s2.shape:
        raise ValueError("volumes of different shapes")
    ks = range(1, s1.shape[0] + 1) if K_range is None else K_range
    return np.array([cor_method(s1[k - 1], s2[k - 1]) for k in ks])


def shell_correlation_ISC(volume1, volume2, K_range=None, ctx=None):
    return shell_correlation(volume1, volume2, K_range, cor_ISC, ctx)


def shell_correlation_FSC(volume1, volume2, K_range=None, ctx=None):
    return shell_correlation(volume1, volume2, K_range, cor_FSC, ctx)


def ISC(density, intensity, K_range=None, ctx=None):
    """Intensity shell correlation of a density against an intensity (spherical_harmonics.jl:417-419)."""
    from .phases import forward
    return shell_correlation_ISC(intensity, absoluteSquare(forward(density, ctx), ctx), K_range, ctx)


def FSC(density, fourier, K_range=None, ctx=None):
    """Fourier shell correlation of a density against a Fourier volume (spherical_harmonics.jl:421-423)."""
    from .phases import forward
    return shell_correlation_FSC(fourier, forward(density, ctx), K_range, ctx)
