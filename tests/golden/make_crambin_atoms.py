"""Writes crambin_atoms.pdb: the ATOM records of the crambin structure that
the reference's structure test loads (tests/test_structure.jl:3,
data/structures/crambin.pdb; PDB entry 3U7T), the only records loadPDB reads.

Run in a container that has the reference checkout:
    python tests/golden/make_crambin_atoms.py /root/reference
"""
import os
import sys

src = os.path.join(sys.argv[1] if len(sys.argv) > 1 else "/root/reference", "data", "structures", "crambin.pdb")
dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "crambin_atoms.pdb")
with open(src) as f:
    lines = [ln for ln in f if ln.startswith("ATOM")]
with open(dst, "w") as f:
    f.writelines(lines)
print(f"{len(lines)} ATOM records -> {dst}")
