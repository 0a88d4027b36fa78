"""libtp's host FP64 special functions (basis precompute) against the golden
vectors: the header is compiled for the host with g++.  CPU only."""
import json
import os
import subprocess

import numpy as np
import pytest

from helpers import seanindex

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "threephotons.jl_amd", "csrc")
GOLD = os.path.join(HERE, "golden")


@pytest.fixture(scope="module")
def driver(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("hm") / "host_math_driver")
    subprocess.run(["g++", "-std=c++17", "-O2", "-I", CSRC, "-o", exe, os.path.join(HERE, "host_math_driver.cpp")],
                   check=True)

    def run(cmds):
        out = subprocess.run([exe], input="\n".join(cmds) + "\n", capture_output=True, text=True, check=True)
        return out.stdout.split("\n")

    return run


def test_wigner3j_golden(driver):
    rows = json.load(open(os.path.join(GOLD, "wigner3j.json")))
    got = driver(["w %d %d %d %d %d %d" % tuple(r[:6]) for r in rows])
    worst = max(abs(float(g) - r[6]) / max(abs(r[6]), 1e-3) for g, r in zip(got, rows))
    assert worst < 1e-14, worst


def test_sphplm_golden(driver):
    rows = json.load(open(os.path.join(GOLD, "sphplm.json")))
    got = driver(["p %d %d %.17g" % (l, m, th) for l, m, th, _ in rows])
    worst = max(abs(float(g) - r[3]) / max(abs(r[3]), 1e-2) for g, r in zip(got, rows))
    assert worst < 1e-13, worst


def test_seanindex(driver):
    cmds, want = [], []
    for L in (1, 7, 25):
        for l in range(L):
            for m in range(-l, l + 1):
                cmds.append(f"s {l} {m} {L}")
                want.append(seanindex(l, m, L))
    got = [int(x) for x in driver(cmds)[: len(want)]]
    assert got == want


def test_gauss_legendre(driver):
    for n in (1, 5, 48, 96):
        rows = driver([f"g {n}"])[:n]
        x = np.array([float(r.split()[0]) for r in rows])
        w = np.array([float(r.split()[1]) for r in rows])
        xr, wr = np.polynomial.legendre.leggauss(n)
        assert np.abs(np.sort(x) - xr).max() < 1e-14
        assert np.abs(w[np.argsort(x)] - wr).max() < 1e-14


def test_fastlog_ulp(driver):
    """The energy epilogues' table-driven log (csrc/tp_fastlog.hpp) against a
    long double reference over 1e-30..1e30, around 1 and at table edges."""
    worst = float(driver(["l 400000"])[0])
    assert worst < 4.0, worst


def test_fastlog_energy_form(driver):
    """The branch-free log of the FP64 energy epilogues (tp_fastlog_e): within
    4 ulp of max(|log x|, 1) of the long double log (an absolute bound, which
    is what the energy sum sees) over 1e-30..1e30 and around 1, and equal to
    tp_fastlog bit for bit away from 1."""
    worst, differ = driver(["e 400000"])[0].split()
    assert float(worst) < 4.0, worst
    assert int(differ) == 0


def test_log_series_ulp(driver):
    """The table-free log of the mixed-precision contraction's energy epilogue
    (tp_log_series) against long double over 1e-30..1e30, around 1 and at
    the sqrt2 fold of its argument reduction."""
    worst = float(driver(["t 400000"])[0])
    assert worst < 4.0, worst


def _bessel(driver, n, xs):
    return [np.array([float(v) for v in ln.split()]) for ln in driver([f"j {n} {x:.17g}" for x in xs])[:len(xs)]]


def test_half_integer_bessel_closed_forms(driver):
    """J_{1/2} = sqrt(2/(pi x)) sin x, J_{3/2} = sqrt(2/(pi x)) (sin x / x - cos x),
    J_{5/2} = sqrt(2/(pi x)) ((3/x^2 - 1) sin x - 3 cos x / x), at the Hankel
    grid's arguments pi k r / K (including the zeros of sin near k r = K n)."""
    xs = sorted({np.pi / K * k * r for K in (5, 26, 38) for k in range(1, K + 1) for r in range(1, K + 1, 3)})
    for x, J in zip(xs, _bessel(driver, 3, xs)):
        f = np.sqrt(2 / (np.pi * x))
        s, c = np.sin(x), np.cos(x)
        want = [f * s, f * (s / x - c), f * ((3 / x ** 2 - 1) * s - 3 * c / x)]
        assert np.allclose(J, want, rtol=0, atol=1e-15 * max(1.0, 3 / x ** 2)), x


def test_half_integer_bessel_high_orders(driver):
    """Orders up to 47.5 against mpmath at 30 digits: within 1e-16 absolute
    and 1e-14 relative wherever |J| > 1e-6."""
    mpmath = pytest.importorskip("mpmath")
    mpmath.mp.dps = 30
    rng = np.random.default_rng(5)
    xs = list(np.pi / 38 * rng.integers(1, 39, 40) * rng.integers(1, 39, 40)) + [0.05, 1e-3, 150.0]
    for x, J in zip(xs, _bessel(driver, 48, xs)):
        for l in range(0, 48, 3):
            t = float(mpmath.besselj(l + 0.5, mpmath.mpf(float(x))))
            assert abs(J[l] - t) <= 1e-16 + (1e-14 * abs(t) if abs(t) > 1e-6 else 0.0), (l, x, J[l], t)
