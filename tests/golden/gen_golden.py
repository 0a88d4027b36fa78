"""Writes the golden fixtures of tests/golden/ from independent libraries.

  wigner3j.json   exact Wigner 3j symbols from sympy.physics.wigner.wigner_3j
                  (the reference calls GSL gsl_sf_coupling_3j,
                  spherical_harmonics.jl:63-65)
  sphplm.json     normalised associated Legendre values with the
                  Condon-Shortley phase from mpmath.legenp at 30 digits
                  (the reference calls GSL gsl_sf_legendre_sphPlm,
                  spherical_harmonics.jl:69-76)
  philox_kat.json Philox4x32-10 outputs from PyTorch's header-only engine
                  (gen_philox_kat.cpp)

Run from the repository root:  python tests/golden/gen_golden.py
"""
import json
import os
import subprocess
import sys

import mpmath as mp
import numpy as np
from sympy.physics.wigner import wigner_3j

HERE = os.path.dirname(os.path.abspath(__file__))


def make_wigner():
    rng = np.random.default_rng(20240601)
    rows = []
    # every symbol with l <= 4 (all parities), selection-rule zeros included
    for l1 in range(5):
        for l2 in range(5):
            for l3 in range(5):
                for m1 in range(-l1, l1 + 1):
                    for m2 in range(-l2, l2 + 1):
                        m3 = -m1 - m2
                        if abs(m3) > l3:
                            continue
                        rows.append((l1, m1, l2, m2, l3, m3))
    # random sample of non-trivial symbols up to l = 22 (even l, as the path uses)
    while len(rows) < 4000:
        l1, l2 = (int(x) for x in rng.integers(0, 12, size=2) * 2)
        lo, hi = abs(l1 - l2), min(l1 + l2, 22)
        l3 = int(rng.integers(lo // 2, hi // 2 + 1)) * 2
        m1 = int(rng.integers(-l1, l1 + 1))
        m2 = int(rng.integers(-l2, l2 + 1))
        m3 = -m1 - m2
        if abs(m3) > l3:
            continue
        rows.append((l1, m1, l2, m2, l3, m3))
    out = []
    for (l1, m1, l2, m2, l3, m3) in rows:
        v = wigner_3j(l1, l2, l3, m1, m2, m3)
        out.append([l1, m1, l2, m2, l3, m3, float(mp.mpf(v.evalf(40)))])
    with open(os.path.join(HERE, "wigner3j.json"), "w") as f:
        json.dump(out, f)
    print("wigner3j.json:", len(out))


def make_sphplm():
    mp.mp.dps = 30
    rng = np.random.default_rng(7)
    out = []
    thetas = [np.pi / 2, 0.1, 1.0, 2.5, np.arccos(0.3)] + list(rng.uniform(0.0, np.pi, size=6))
    for th in thetas:
        x = mp.cos(mp.mpf(th))
        for l in range(0, 23):
            for m in range(-l, l + 1):
                am = abs(m)
                v = mp.sqrt((2 * l + 1) / (4 * mp.pi) * mp.factorial(l - am) / mp.factorial(l + am)) * mp.legenp(l, am, x)
                if m < 0 and am % 2 == 1:
                    v = -v
                out.append([l, m, float(th), float(v)])
    with open(os.path.join(HERE, "sphplm.json"), "w") as f:
        json.dump(out, f)
    print("sphplm.json:", len(out))


def make_philox_kat():
    import torch
    inc = os.path.join(os.path.dirname(torch.__file__), "include")
    exe = "/tmp/gen_philox_kat"
    subprocess.run(["g++", "-std=c++17", "-O1", "-I" + inc, "-o", exe, os.path.join(HERE, "gen_philox_kat.cpp")],
                   check=True)
    with open(os.path.join(HERE, "philox_kat.json"), "w") as f:
        subprocess.run([exe], stdout=f, check=True)
    print("philox_kat.json written")


if __name__ == "__main__":
    which = sys.argv[1:] or ["wigner", "sphplm", "philox"]
    if "wigner" in which:
        make_wigner()
    if "sphplm" in which:
        make_sphplm()
    if "philox" in which:
        make_philox_kat()
