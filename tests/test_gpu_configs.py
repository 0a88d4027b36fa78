"""Parity at the benchmark workloads themselves (L=18, LMAX=25, N=32):

  full size   1000 point atoms, K=26 (3,276 k-triples x 2,048 bins), the
              1e9-count Poisson histogram, start = randomStartStructure(21):
              c3 against the factored oracle on every column, per column;
              the energy against the oracle's; 12 annealing steps against the
              oracle's driver with identical decisions.
  batching    20 runs (seeds 1000..1019, K_r ~ U{10..26} from the seed,
              randomStartStructure(seed), the histogram cut to K_r shells),
              64 proposals each, FP16 hi/lo: all 1,280 energies of one batched
              evaluation against single FP64 evaluations of the same rotated
              structures, two of them against the oracle, and every run's
              energy after 3 graph-replayed sweeps against an FP64 evaluation
              of its coefficients.
  learned     the same batch with 48 of 64 slots from a trained network.

Bars: FP64 c3 and energies 1e-10 relative (c3 per k-triple column:
max_r |d| / max_r |ref|); the batched mixed-precision energies 1e-6 relative
to FP64 (their c3 is held to 1e-5 by test_gpu_precision)."""
import math
import sys

import numpy as np
import pytest

from helpers import c2_invariants, random_quats

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
C3_TOL = 1e-10
MIXED_E_TOL = 1e-6
SEEDS = tuple(range(1000, 1020))
NPROP = 64


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def full():
    import tp_amd
    from tp_amd import workload
    tp_amd.HIP_init()
    cfg = workload.CONFIGS["full"]
    say("full size: intensity, basis, histogram")
    atoms = workload.atoms_in_ball(cfg["natoms"], cfg["radius"], cfg["atoms_seed"])
    truth = tp_amd.intensity_from_atoms(atoms, cfg["KMAX"], cfg["qmax"], cfg["LMAX"], cfg["L"], cfg["ntheta"],
                                        cfg["nphi"])
    basis = tp_amd.HIP_calculate_basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], 0.0, truth.dq())
    c3 = tp_amd.FullCorrelation_parallized(truth, basis, return_raw=True)
    c3ref, _ = workload.poisson_histogram(c3, cfg["N"], cfg["K"], cfg["counts"], cfg["hist_seed"])
    start = workload.random_start_structure(truth, cfg["L"], cfg["start_seed"])
    return dict(tp=tp_amd, workload=workload, cfg=cfg, truth=truth, basis=basis, c3ref=c3ref, start=start)


def test_full_c3_every_column(full, oracle):
    tp, cfg, start, basis = full["tp"], full["cfg"], full["start"], full["basis"]
    say("full size: c3 of the start structure, all 3,276 columns")
    got = tp.FullCorrelation_parallized(start, basis, return_raw=True)
    ob = oracle.Basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], 0.0, start.dq())
    want = ob.c3_factored(start.coeff[:cfg["K"]])
    assert got.shape == want.shape == (2048, 3276)
    col = np.abs(got - want).max(axis=0) / np.abs(want).max(axis=0)
    say(f"  per-column error max {col.max():.2e}, median {np.median(col):.2e}")
    assert col.max() < C3_TOL
    # and a spread of columns against the dense reference formulation (B . PA)
    for lo in (0, 1111, 2222, 3270):
        d = ob.c3_dense(start.coeff[:cfg["K"]], lo, lo + 6)
        assert (np.abs(got[:, lo:lo + 6] - d).max(axis=0) / np.abs(d).max(axis=0)).max() < C3_TOL


def test_full_energy_on_poisson_histogram(full, oracle):
    tp, cfg, start, basis, c3ref = full["tp"], full["cfg"], full["start"], full["basis"], full["c3ref"]
    say("full size: energy on the 1e9-count histogram")
    ob = oracle.Basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], 0.0, start.dq())
    c3o = ob.c3_factored(start.coeff[:cfg["K"]])
    Eo = oracle.energy(c3o, cfg["N"], c3ref, cfg["K"], 1, cfg["K"])
    E = tp.energy(start, basis, c3ref)
    st = tp.DeviceState(start, basis)
    Es = st.energy(c3ref)
    say(f"  E {E!r} oracle {Eo!r}")
    assert abs(E - Eo) / abs(Eo) < E_TOL and E == Es
    # the walk from K3_range 2:26 (correlations.jl:211-216)
    E2 = tp.energy(start, basis, c3ref, range(2, cfg["K"] + 1))
    E2o = oracle.energy(c3o, cfg["N"], c3ref, cfg["K"], 2, cfg["K"])
    assert abs(E2 - E2o) / abs(E2o) < E_TOL


def test_full_anneal_steps_match_oracle(full, oracle):
    """12 steps of the plain annealing at the full size (initial stepsize pi/4,
    decay 0.99998, factor 1.01, T factor 0.1; 4 temperature probes), device
    against the oracle's driver on the same Philox streams."""
    tp, cfg, start, basis, c3ref = full["tp"], full["cfg"], full["start"], full["basis"], full["c3ref"]
    nsteps, ntemp = 12, 4
    say("full size: oracle annealing")
    ob = oracle.Basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], 0.0, start.dq())
    ro = oracle.anneal(ob, c3ref, start.coeff, nsteps, math.pi / 4, 1.01, 0.1, 0.99998, seed=7, ntemp=ntemp)
    say("full size: device annealing")
    st = tp.DeviceState(start, basis)
    rg = tp.rotate_all_at_once(st, basis, c3ref, nsteps, math.pi / 4, 1.01, 0.1, 0.99998, seed=7, ntemp=ntemp)
    say(f"  accepted {rg['accepted']} oracle {ro['accepted']}")
    assert rg["E0"] == pytest.approx(ro["E0"], rel=E_TOL)
    assert rg["T0"] == pytest.approx(ro["T0"], rel=1e-8)
    assert np.array_equal(rg["accepted"], ro["accepted"])
    assert 0 < rg["accepted"].sum() < nsteps
    assert np.abs(rg["E_new"] - ro["E_new"]).max() < E_TOL * abs(ro["E0"])
    final = st.read()
    assert np.abs(final - ro["coeff"]).max() < 1e-10 * np.abs(ro["coeff"]).max()


def _batch(full, nlearn=0, mlp=None, prec="fp16x3"):
    tp, wl, cfg, truth, basis, c3ref = (full[k] for k in ("tp", "workload", "cfg", "truth", "basis", "c3ref"))
    Ks = [int(np.random.default_rng(s).integers(10, 27)) for s in SEEDS]
    starts = [wl.random_start_structure(truth, cfg["L"], s) for s in SEEDS]
    refs = [np.ascontiguousarray(c3ref[:, :, :K, :K, :K]) for K in Ks]
    basis.set_precision(prec)
    B = tp.AnnealBatch(basis, refs, starts, SEEDS, nprop=NPROP, hist_cap=16)
    return Ks, starts, refs, B


def _fp64_energy(full, c, K, ref):
    """FP64 energy of run coefficients c (>= K shells) on the K=26 basis, walking
    K3_range 1:K (the first klength(K) columns are the k-triples of K shells)."""
    tp, basis, cfg = full["tp"], full["basis"], full["cfg"]
    cc = np.zeros((cfg["K"], c.shape[1]), np.complex128)
    cc[:K] = c[:K]
    return tp.energy(cc, basis, ref, range(1, K + 1))


@pytest.mark.parametrize("prec", ["fp16x3", "fp16x3_t64"])
def test_batched_evaluation_matches_fp64(full, oracle, prec):
    """fp16x3: compensated FP16 triple products; fp16x3_t64: FP64 ones on the
    sorted l-triples (runs of K > 16 take both shell halves)."""
    tp, cfg, basis = full["tp"], full["cfg"], full["basis"]
    say(f"batch ({prec}): 20 runs x 64 proposals")
    Ks, starts, refs, B = _batch(full, prec=prec)
    assert sorted(set(Ks)) == sorted(set(Ks)) and min(Ks) >= 10 and max(Ks) <= 26
    rng = np.random.default_rng(2024)
    q = np.stack([random_quats(rng, cfg["L"] // 2) for _ in range(len(Ks) * NPROP)])
    E = B.eval_quats(q)
    assert E.shape == (len(Ks) * NPROP,) and np.all(np.isfinite(E))
    basis.set_precision("fp64")
    worst = 0.0
    for g, K in enumerate(Ks):
        for p in range(NPROP):
            e = g * NPROP + p
            c = tp.rotate(starts[g], basis, q[e])
            Es = _fp64_energy(full, c, K, refs[g])
            worst = max(worst, abs(E[e] - Es) / abs(Es))
    say(f"  1,280 batched energies vs FP64: max rel {worst:.2e}")
    assert worst < MIXED_E_TOL
    # two slots through the oracle: the run with the fewest and the one with the most shells
    for g in (int(np.argmin(Ks)), int(np.argmax(Ks))):
        K, e = Ks[g], g * NPROP + 5
        c = tp.rotate(starts[g], basis, q[e])
        ob = oracle.Basis(cfg["L"], cfg["LMAX"], cfg["N"], K, 0.0, starts[g].dq())
        Eo = oracle.energy(ob.c3_factored(c[:K]), cfg["N"], refs[g], K, 1, K)
        say(f"  run {g} (K={K}) slot 5: batch {E[e]!r} oracle {Eo!r}")
        assert abs(E[e] - Eo) / abs(Eo) < MIXED_E_TOL


def _check_runs_after_sweeps(full, Ks, starts, refs, B, nsw):
    tp, cfg, basis = full["tp"], full["cfg"], full["basis"]
    basis.set_precision("fp64")
    hist = B.history(nsw)
    for g, K in enumerate(Ks):
        st = B.group(g)
        assert st["nsweep"] == nsw
        assert hist[-1, g] == st["E"]
        c = B.coeff(g)
        Es = _fp64_energy(full, c, K, refs[g])
        assert abs(st["E"] - Es) / abs(Es) < MIXED_E_TOL, (g, st["E"], Es)
        a, b2 = c2_invariants(starts[g].coeff, K, cfg["L"], cfg["LMAX"]), c2_invariants(c, K, cfg["L"], cfg["LMAX"])
        assert np.abs(a - b2).max() < 1e-11 * np.abs(a).max()


def test_batched_sweeps_keep_exact_energies(full):
    say("batch: probe + 3 sweeps (graph replays)")
    Ks, starts, refs, B = _batch(full)
    B.probe(0.1)
    B.run(3, factor=1.05)
    _check_runs_after_sweeps(full, Ks, starts, refs, B, 3)


def test_learned_slots_at_full_size(full):
    """48 of 64 slots from a network trained on the batch's own logged moves."""
    tp = full["tp"]
    say("learned: log, train, 3 sweeps with 48 learned slots")
    Ks, starts, refs, B = _batch(full)
    B.log(400000)
    B.probe(0.1)
    B.run(4, factor=1.05)
    n = B.log_count()
    assert n > 0
    m = tp.MlpProposer(seed=3)
    tr = m.train_log(B, steps=60, batch=512, lr=1e-3, seed=4)
    assert np.all(np.isfinite(tr))
    Ks2, starts2, refs2, B2 = _batch(full)
    B2.set_proposer(m, 48, noise=0.25)
    B2.probe(0.1)
    B2.run(3, factor=1.05)
    _check_runs_after_sweeps(full, Ks2, starts2, refs2, B2, 3)
