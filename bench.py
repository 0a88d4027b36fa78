"""Benchmark of the plain FP64 c3 + Bayes-energy path on one MI355X.

Workload (the full-size plain configuration): 1000 seeded random point atoms
in a ball of radius 26, KMAX=38 shells (qmax = pi), intensity expanded to L=18
(even l), c3 over all 3276 k1>=k2>=k3 triples of K=26 shells on the N=32
alpha/beta grid (2048 bins), scored against a Poisson histogram of 1e9 counts.

A step is one c3 + energy evaluation of the device-resident structure
(tp_state_energy: triple products, FP64-MFMA contraction with the fused
energy epilogue, final reduction).  value = evaluations per second.
The 2000-step plain annealing run (tp_anneal) gives MC steps/s.

    python bench.py [--steps K] [--warmup W] [--anneal-steps S] [--no-cpu]
"""
import argparse
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "threephotons.jl_amd"))

METRIC = "c3 evals/s + MC annealing steps/s at K=26 L=18 N=32; time-to-energy vs plain FP64"
SWEEP_FACTOR = 1.05   # angle-bound growth/shrink per sweep of the batched annealing (a sweep stands for many steps)
FP64_PEAK_TFLOPS = 78.6     # MI355X dense FP64 (matrix and vector)
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def folded_classes(L):
    """Even (m2, mu3) pairs with |m2|,|mu3|,|m2+mu3| <= L, (m2,mu3) ~ (-m2,-mu3)."""
    n = 0
    for m2 in range(0, L + 1, 2):
        for mu3 in range(-L, L + 1, 2):
            if (m2 == 0 and mu3 < 0) or abs(m2 + mu3) > L:
                continue
            n += 1
    return n


def sep_grid_flops(L, N):
    """FP64 flops per k-triple column of k_c3sep on the folded grid (lambda = 0,
    even N: the N^2/2 orbit representatives i < N/2, j < N): stage A (C, S over
    mu3 at the N values of beta) 4 n_m N, stage B (over m2 = 0, 2, .., L at
    each representative bin) 4 n_m2 N^2/2."""
    n_m, n_m2 = folded_classes(L), L // 2 + 1
    return 4.0 * n_m * N + 2.0 * n_m2 * N * N


def lt_n3_sum(L):
    """Sum over l-triples (even l, triangle rule) of 2 l3 + 1."""
    s = 0
    for l1 in range(0, L + 1, 2):
        for l2 in range(0, L + 1, 2):
            for l3 in range(abs(l1 - l2), min(l1 + l2, L) + 1, 2):
                s += 2 * l3 + 1
    return s


def sorted_lt_flops(L, K):
    """Algorithmic FLOP of the triple products on sorted l-triples (k_triple_sym)
    for K shells: per sorted l-triple s = (L1 >= L2 >= L3), stage 1 forms T for
    all K^2 ordered shell pairs over its E_s entries (one complex MAC, 8 flop,
    each) and stage 2 the real part of the mu3 sum for all K^3 ordered shell
    triples (2 real MACs, 4 flop, per mu3 value); E_s counts the (m1, m2) pairs
    allowed by the selection rules (|m1 + m2| <= L3)."""
    e_sum = n3_sum = 0
    for l1 in range(0, L + 1, 2):
        for l2 in range(0, l1 + 1, 2):
            for l3 in range(0, l2 + 1, 2):
                if l3 < l1 - l2:
                    continue
                n3_sum += 2 * l3 + 1
                e_sum += sum(1 for m1 in range(-l1, l1 + 1) for m2 in range(-l2, l2 + 1) if abs(m1 + m2) <= l3)
    return 8.0 * K * K * e_sum + 4.0 * K ** 3 * n3_sum


# HBM bytes per launch of the plain path's dominant kernel (k_triple_sym) at
# the full-size workload, from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per
# dispatch), scaled by what the same counters read for a 256 MiB device copy
# (FETCH_SIZE 0.5x, WRITE_SIZE 1.0x of the bytes moved); the summary with the
# commands is profiles/hbm_counters_full_L18_K26_N32.md.
TRIPLE_FETCH_KIB, TRIPLE_WRITE_KIB = 1683.8, 13708.3
FETCH_SCALE, WRITE_SCALE = 2.0, 1.0


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def stage_times(ctx, state, c3ref, n):
    """Per-stage device times (ms) of an evaluation, HIP events on the library
    stream, in a pass of its own (the timed loop runs without instrumentation)."""
    ctx.set_timing(True)
    kern = np.zeros(4)
    for _ in range(n):
        state.energy(c3ref)
        kern += ctx.last_timings()
    ctx.set_timing(False)
    return kern / n


def setup(cfg, tp, workload):
    t0 = time.time()
    atoms = workload.atoms_in_ball(cfg["natoms"], cfg["radius"], cfg["atoms_seed"])
    truth = tp.intensity_from_atoms(atoms, cfg["KMAX"], cfg["qmax"], cfg["LMAX"], cfg["L"], cfg["ntheta"], cfg["nphi"])
    log(f"intensity: {time.time() - t0:.2f}s")
    t0 = time.time()
    basis = tp.HIP_calculate_basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], 0.0, truth.dq())
    log(f"basis (basislen {basis.basislen}, n_lt {basis.n_lt}, klength {basis.klength}): {time.time() - t0:.2f}s")
    c3 = tp.FullCorrelation_parallized(truth, basis, return_raw=True)
    c3ref, _ = workload.poisson_histogram(c3, cfg["N"], cfg["K"], cfg["counts"], cfg["hist_seed"])
    start = workload.random_start_structure(truth, cfg["L"], cfg["start_seed"])
    return truth, basis, c3ref, start


def cpu_baseline(cfg, start, c3ref, ncols, nthreads):
    """The oracle (C, FP64, OpenMP over `nthreads` host threads) in the two forms:
    the reference formulation (dense basislen-inner contraction B*PA with the
    triple-product stage, correlations.jl:157-193) on a sample of k-triple
    columns, scaled to whole evaluations; the factorised formulation (the GPU's
    algorithm) on three whole evaluations; the MC move; and the small
    configuration in full."""
    from oracle import pyoracle as po
    from tp_amd import workload
    ob = po.Basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"])
    t0 = time.perf_counter()
    ob.c3_dense(start.coeff, 0, ncols, nthreads=nthreads)
    t_dense = time.perf_counter() - t0
    dense_evals = (ncols / ob.klength) / t_dense
    t_fact = []
    for _ in range(3):
        t0 = time.perf_counter()
        c3f = ob.c3_factored(start.coeff, nthreads=nthreads)
        po.energy(c3f, cfg["N"], c3ref, cfg["K"], 1, cfg["K"])
        t_fact.append(time.perf_counter() - t0)
    steps = np.array([np.pi / 4 * (l - 1) for l in range(2, cfg["L"] + 1, 2)])
    t0 = time.perf_counter()
    for i in range(20):
        po.mc_propose(1000, i + 1, 0x4D4F5645, cfg["L"], steps, start.coeff, cfg["LMAX"])
    t_move = (time.perf_counter() - t0) / 20
    # the small configuration, both forms in full
    sm = workload.CONFIGS["small"]
    atoms = workload.atoms_in_ball(sm["natoms"], sm["radius"], sm["atoms_seed"])
    coeff0 = po.intensity_from_atoms(atoms, sm["KMAX"], sm["qmax"] / sm["KMAX"], sm["LMAX"], sm["L"], sm["ntheta"],
                                     sm["nphi"], nthreads=nthreads)
    ob0 = po.Basis(sm["L"], sm["LMAX"], sm["N"], sm["K"])
    t0 = time.perf_counter()
    c3d0 = ob0.c3_dense(coeff0, nthreads=nthreads)
    t_d0 = time.perf_counter() - t0
    ref0, _ = workload.poisson_histogram(c3d0, sm["N"], sm["K"], sm["counts"], sm["hist_seed"])
    t0 = time.perf_counter()
    c3f0 = ob0.c3_factored(coeff0, nthreads=nthreads)
    E0 = po.energy(c3f0, sm["N"], ref0, sm["K"], 1, sm["K"])
    t_f0 = time.perf_counter() - t0
    # photon-triplet histograms: the oracle on 1024 pictures of the histogram workload
    hc = HIST_CFG
    ph, off = workload.beamstop_pictures(1024, hc["nphot"], hc["qcut"], hc["seed"], beamstop_width=hc["qcut"] / 20.0)
    t0 = time.perf_counter()
    po.hist_pictures(ph, off, hc["qcut"] / hc["K2"], hc["N"], hc["K2"], hc["K3"])
    t_hist = time.perf_counter() - t0
    # two-photon correlation and retrieveSolution (numpy oracle, 38 shells of the start structure)
    from oracle import c2_oracle
    t0 = time.perf_counter()
    c2s = c2_oracle.two_photons(start.coeff, cfg["L"], cfg["LMAX"], cfg["N"], start.KMAX)
    t_c2 = time.perf_counter() - t0
    t0 = time.perf_counter()
    c2_oracle.retrieve_solution(c2s, cfg["L"], cfg["LMAX"], 1, start.KMAX)
    t_ret = time.perf_counter() - t0
    # rotational fit: the oracle's similarity of two rotated copies (38 shells, L = 18, LMAX = 25)
    from oracle import fit_oracle
    t0 = time.perf_counter()
    for e in ((0.3, 0.2, 0.1), (1.0, 2.0, 0.5)):
        rs = fit_oracle.rotate_structure(start.coeff, *e, start.KMAX, range(2, cfg["L"] + 1, 2), cfg["LMAX"])
        fit_oracle.similarity(start.coeff, rs, cfg["LMAX"], 1, start.KMAX)
    t_fit = (time.perf_counter() - t0) / 2
    # phase retrieval: the oracle's sRAAR iterations on the 38-shell start structure (LMAX = 25)
    from oracle import phase_oracle
    php = np.random.default_rng(1).uniform(0, 2 * np.pi, (start.KMAX, 4 * cfg["LMAX"] ** 2))
    t0 = time.perf_counter()
    phase_oracle.sraar(start.coeff, cfg["LMAX"], 1, 0.75, 0.9, 350.0, 0.3, php)
    t_ph1 = time.perf_counter() - t0
    t0 = time.perf_counter()
    phase_oracle.sraar(start.coeff, cfg["LMAX"], 3, 0.75, 0.9, 350.0, 0.3, php)
    t_ph = (time.perf_counter() - t0 - t_ph1) / 2
    # picture synthesis: the oracle on 8 pictures of the same workload
    from oracle import picture_oracle
    surf = picture_oracle.surfaces(start.coeff, cfg["LMAX"])
    t0 = time.perf_counter()
    for p in range(8):
        picture_oracle.points_per_orientation(surf, cfg["LMAX"], start.KMAX, start.rmax, start.rmax, start.rmax / 3.0,
                                              1000, 0.0, 0.0, 7, p)
    t_pic = (time.perf_counter() - t0) / 8
    # structure: the oracle's density sampling of crambin on 4 of the 38 shells (LMAX = 35), scaled to 38
    from oracle import structure_oracle
    from tp_amd import structure
    sat = structure.atom_table(structure.loadPDB(STRUCTURE_PDB))
    t0 = time.perf_counter()
    structure_oracle.density_surface(sat, 35, 4, 4.0)
    t_dens = (time.perf_counter() - t0) * 38 / 4
    tf = float(np.median(t_fact))
    return dict(value=dense_evals, unit="c3+energy evals/s", cores=nthreads, kind="port",
                sample=f"dense reference formulation on {ncols} of {ob.klength} k-triple columns "
                       f"({t_dense:.1f}s), scaled to whole evaluations; host '{cpu_model()}'",
                factored_evals_per_s=1.0 / tf,
                factored_sample=f"3 full factorised evaluations (F.As + energy), median {tf:.2f}s",
                mc_steps_per_s_factored=1.0 / (tf + t_move), move_seconds=t_move,
                histogram_pictures_per_s=1024 / t_hist,
                histogram_sample="or_hist_pictures (1 thread) on 1024 pictures of the histogram workload",
                two_photons_s=t_c2, retrieve_s=t_ret,
                two_photon_sample="numpy oracle: twoPhotons of the 38-shell start structure and retrieveSolution of it",
                fit_rotations_per_s=1.0 / t_fit,
                fit_sample="numpy oracle: rotateStructure + similarity of 2 rotations, 38 shells, L=18, LMAX=25",
                sraar_ms_per_iteration=1e3 * t_ph,
                picture_pictures_per_s=1.0 / t_pic,
                picture_sample="numpy oracle: pointsPerOrientation of 8 pictures, 1000 incident photons each",
                structure_density_s=t_dens,
                structure_sample="numpy oracle: crambin density sampling (670 atoms, LMAX=35) on 4 shells, x 38/4",
                sraar_sample="numpy oracle: sRAAR with 3 and with 1 iterations, 38 shells, LMAX=25; the difference / 2",
                small_config={"dense_evals_per_s": 1.0 / t_d0, "factored_evals_per_s": 1.0 / t_f0,
                              "energy": E0, "klength": ob0.klength, "R": 2 * sm["N"] ** 2})


def precision_study(n):
    """Mixed-precision c3 and energy against the FP64 path over n random
    orientations (independent Haar SO(2l+1) per even l, seeds 0..n-1,
    randomStartStructure, determination.jl:4-17) of the full-size structure.
    c3 errors: per k-triple slice max_r|dc|/max_r|c3| (max over slices), the
    RMS over all bins of |dc|/max_r|c3| per slice, and the normwise
    ||dc||_2/||c3||_2."""
    import tp_amd as tp
    from tp_amd import workload
    cfg = workload.CONFIGS["full"]
    tp.HIP_init()
    truth, basis, c3ref, _ = setup(cfg, tp, workload)
    c3ref = basis.histogram(c3ref)
    out = {"metric": "mixed-precision c3 / energy error vs FP64", "orientations": n, "config": "full (L=18,K=26,N=32)"}
    modes = ("fp16x3", "fp16x3_t64", "triple16", "bf16")
    acc = {m: {"slice_max": [], "rms": [], "norm2": [], "dE_abs": [], "dE_rel": []} for m in modes}
    for seed in range(n):
        st = workload.random_start_structure(truth, cfg["L"], seed)
        dev = tp.DeviceState(st, basis)
        basis.set_precision("fp64")
        c64 = dev.c3()
        E64 = dev.energy(c3ref)
        scale = np.abs(c64).max(axis=0)
        for m in modes:
            basis.set_precision(m)
            c = dev.c3()
            E = dev.energy(c3ref)
            d = np.abs(c - c64) / scale
            a = acc[m]
            a["slice_max"].append(float(d.max()))
            a["rms"].append(float(np.sqrt((d * d).mean())))
            a["norm2"].append(float(np.linalg.norm(c - c64) / np.linalg.norm(c64)))
            a["dE_abs"].append(abs(E - E64))
            a["dE_rel"].append(abs(E - E64) / abs(E64))
        basis.set_precision("fp64")
        log(f"orientation {seed}: " + "  ".join(f"{m} {acc[m]['slice_max'][-1]:.2e}" for m in modes))
    for m, a in acc.items():
        out[m] = {"c3_slice_rel_err_max": max(a["slice_max"]), "c3_slice_rel_err_rms": float(np.sqrt(np.mean(np.square(a["rms"])))),
                  "c3_norm2_rel_err_max": max(a["norm2"]), "energy_abs_err_max": max(a["dE_abs"]),
                  "energy_rel_err_max": max(a["dE_rel"])}
    print(json.dumps(out), flush=True)


def gradient_profile(ncalls, ncand=16, reps=10):
    """Instead of the benchmark: the parts of a gradient move on the full-size
    workload, timed from the host (each after one untimed call), then ncalls
    more gradient calls (the kernel-trace target of the gradient profile)."""
    import tp_amd as tp
    from tp_amd import descent, workload
    cfg = workload.CONFIGS["full"]
    tp.HIP_init()
    truth, basis, c3ref, start = setup(cfg, tp, workload)
    c3ref = basis.histogram(c3ref)
    basis.set_precision("fp16x3")
    c = start.coeff.copy()
    gm = descent.GradientMoves(basis)

    def timed(f):
        r = f()
        t0 = time.perf_counter()
        for _ in range(reps):
            r = f()
        return 1e3 * (time.perf_counter() - t0) / reps, r

    t_grad, (_, g) = timed(lambda: tp.energy_grad(c, basis, c3ref))
    t_riem, om = timed(lambda: gm.riemannian(c, g))
    etas = np.geomspace(0.1, 10.0, ncand) / gm.norm(om)
    t_cand, cands = timed(lambda: gm.candidates(c, om, etas, basis.K))
    t_batch, _ = timed(lambda: tp.energy_batch(list(cands), [basis] * ncand, [c3ref] * ncand))
    t_e64, _ = timed(lambda: descent._energy64(cands[0], basis, c3ref, None))
    for _ in range(ncalls):
        tp.energy_grad(c, basis, c3ref)
    print(json.dumps({"metric": "gradient move parts (ms, host-timed)", "energy_grad": t_grad, "riemannian": t_riem,
                      "candidates_x%d" % ncand: t_cand, "energy_batch_fp16x3_x%d" % ncand: t_batch,
                      "energy_fp64": t_e64}), flush=True)


def batch_leg(cfg, tp, workload, truth, basis, c3ref, nsweeps, nprop=64, seeds=tuple(range(1000, 1020)),
              prec="fp16x3"):
    """Concurrent annealing runs scored together: run r (seed 1000+r) has
    K_r ~ U{10..26} shells (drawn from its seed), starts from
    randomStartStructure(seed) and scores against the full histogram cut to
    K_r shells; every sweep evaluates nprop rotation proposals per run (one
    graph replay for all runs) on the sorted l-triples' mixed-precision
    triple products (FP32 stage 1, FP16 hi/lo stage 2), then the dense FP16
    hi/lo contraction c3 = F.As (prec fp16x3) or the factored one, G = H.As on
    FP16 hi/lo MFMA and the FP64 separated grid (fp16x3_t64)."""
    Ks = [int(np.random.default_rng(s).integers(10, 27)) for s in seeds]
    starts = [workload.random_start_structure(truth, cfg["L"], s) for s in seeds]
    refs = [np.ascontiguousarray(c3ref[:, :, :K, :K, :K]) for K in Ks]
    basis.set_precision(prec)
    t0 = time.perf_counter()
    B = tp.AnnealBatch(basis, refs, starts, seeds, nprop=nprop, hist_cap=nsweeps + 8)
    t_create = time.perf_counter() - t0
    t0 = time.perf_counter()
    tune = B.autotune(reps=3)
    t_tune = time.perf_counter() - t0
    B.probe(0.1)
    E0 = [B.group(g)["E"] for g in range(len(Ks))]
    B.run(2, factor=SWEEP_FACTOR)   # graph capture and first replays, untimed
    t0 = time.perf_counter()
    B.run(nsweeps, factor=SWEEP_FACTOR)
    el = time.perf_counter() - t0
    st = [B.group(g) for g in range(len(Ks))]
    cols = sum(K * (K + 1) * (K + 2) // 6 for K in Ks) * nprop
    ne = len(Ks) * nprop
    ms_sweep = 1e3 * el / nsweeps
    if prec == "fp16x3_t64":
        # factored contraction: G = H.As, then the separated grid (stages A and B)
        n_m = folded_classes(cfg["L"])
        flops = (2.0 * n_m * basis.n_lt + sep_grid_flops(cfg["L"], cfg["N"])) * cols
        contraction = "k_hgemm_g (G = H.As, FP16 hi/lo MFMA) + k_c3sep (c3 = Bm.G, FP64 separated grid, energy)"
    else:
        flops = 2.0 * basis.R * basis.n_lt * cols
        contraction = "k_split_batch + k_gemm16 (c3 = F.As, FP16 hi/lo MFMA, energy)"
    tile_ms = tune["tile_ms"][tune["candidate"]]
    basis.set_precision("fp64")
    log(f"batch ({prec}): {len(Ks)} runs K={Ks}, {ne} evals ({cols} k-triple columns) per sweep: "
        f"{ms_sweep:.2f} ms/sweep, {ne / (ms_sweep * 1e-3):.0f} evals/s; tune {tune}")
    return {
        "runs": len(Ks), "K": Ks, "proposals_per_run": nprop, "evals_per_sweep": ne,
        "kt_columns_per_sweep": cols, "sweeps": nsweeps, "ms_per_sweep": ms_sweep,
        "sweeps_per_s": nsweeps / el, "evals_per_s": ne * nsweeps / el,
        "kt_columns_per_s": cols * nsweeps / el,
        "autotune": tune, "setup_s": t_create, "autotune_s": t_tune,
        "contraction": contraction, "contraction_ms": tile_ms,
        "contraction_tflops_algorithmic": flops / (tile_ms * 1e-3) / 1e12,
        "triple_ms": min(tune["order_ms"]),
        "E0": E0, "E_final": [s["E"] for s in st], "accepted": [s["nacc"] for s in st],
        "precision": prec,
        "mode": "%s: fp16x3 contraction on %s triple products, FP64 rotations/energy; random slots = the "
                "reference's plane rotation in a Haar frame (|b| <= pi/4); best-of-nprop + Metropolis per sweep; "
                "angle factor %.2f/sweep" % (prec, "sorted l-triple (FP32 stage 1, FP16 hi/lo stage 2)",
                                             SWEEP_FACTOR),
    }


def _time_to(Es, t_of, target):
    """First time at which the energy trace reaches target (None if never)."""
    idx = np.nonzero(np.asarray(Es) <= target)[0]
    return None if idx.size == 0 else float(t_of(int(idx[0])))


STRUCTURE_PDB = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tests", "golden", "crambin_atoms.pdb")
HIST_CFG = dict(N=32, K2=38, K3=26, qcut=38.0, nphot=100, seed=5)   # beamstop-only pictures of run_generate_histograms.jl


def hist_leg(tp, workload, npic):
    """Photon-pair / photon-triplet histograms (tp_hist_*) of npic synthetic
    pictures: 100 uniform incident points of [-qcut, qcut]^2 per picture, the
    beamstop (qcut/20) and |p| > qcut dropped (beamstop_shot, datagen.jl:146-160;
    the beamstop-only runs of jobs/run_generate_histograms.jl:327), N = 32,
    K2 = 38, K3 = 26, dq = qcut / K2 = 1.  Device time of the histogram kernel
    (HIP events; photons already on the device) and the call including the
    photon upload."""
    c = HIST_CFG
    ph, off = workload.beamstop_pictures(npic, c["nphot"], c["qcut"], c["seed"], beamstop_width=c["qcut"] / 20.0)
    ctx = tp.HIP_init()
    h = tp.PhotonHistogram(c["N"], c["K2"], c["K3"], c["qcut"] / c["K2"])
    h.add(ph[:off[64]], off[:65])   # warm-up
    h = tp.PhotonHistogram(c["N"], c["K2"], c["K3"], c["qcut"] / c["K2"])
    ctx.set_timing(True)
    t0 = time.perf_counter()
    h.add(ph, off)
    wall = time.perf_counter() - t0
    kern_ms = float(ctx.last_timings()[0])
    ctx.set_timing(False)
    r = h.read()
    n1, n2, n3 = int(r["n1"].sum()), int(r["n2"].sum()), int(r["n3"].sum())
    out = {"workload": "%d pictures x %d incident points (beamstop-only), N=32, K2=38, K3=26" % (npic, c["nphot"]),
           "photons": n1, "pairs": n2, "triplets": n3, "kernel_ms": kern_ms, "call_s": wall,
           "pictures_per_s": npic / (kern_ms * 1e-3), "triplets_per_s": n3 / (kern_ms * 1e-3),
           "pictures_per_s_with_upload": npic / wall}
    log(f"histograms: {npic} pictures ({n1} photons, {n3} triplets) in {kern_ms:.2f} ms of kernel "
        f"({out['pictures_per_s']:.3g} pictures/s, {out['triplets_per_s']:.3g} triplets/s); call {1e3 * wall:.1f} ms")
    return out


def c2_leg(tp, truth, basis, reps):
    """Two-photon correlation of the full structure (twoPhotons, correlations.jl:229-243:
    38 shells, L = 18, N = 32) on the device, and retrieveSolution
    (correlations.jl:262-319, lambda = 0, host) of that c2 over shells 1..38.
    Kernel time by HIP events; the call includes the coefficient upload and the
    c2 download.  The round-trip residual |c2(retrieved) - c2| / max|c2| is
    reported: it is not 0 for this structure, which has l > 2k - 2 content on
    its inner shells that the reference's fit leaves out by design."""
    K2 = truth.KMAX
    ctx = tp.HIP_init()
    c2 = tp.twoPhotons(truth, basis, K2)
    ctx.set_timing(True)
    ks = []
    t0 = time.perf_counter()
    for _ in range(reps):
        tp.twoPhotons(truth, basis, K2)
        ks.append(ctx.last_timings()[0])
    call = (time.perf_counter() - t0) / reps
    ctx.set_timing(False)
    t0 = time.perf_counter()
    vol, info = tp.retrieveSolution(c2, basis.L, truth.LMAX, K2, truth.rmax, return_info=True)
    t_ret = time.perf_counter() - t0
    back = tp.twoPhotons(vol, basis, K2)
    resid = float(np.abs(back - c2).max() / np.abs(c2).max())
    kern = float(np.mean(ks))
    out_bytes = 8.0 * basis.N * K2 * K2
    out = {"workload": f"twoPhotons of {K2} shells, L={basis.L}, N={basis.N}; retrieveSolution over 1:{K2}",
           "kernel_ms": kern, "call_ms": 1e3 * call, "c2_write_GBps": out_bytes / (kern * 1e-3) / 1e9,
           "retrieve_s": t_ret, "retrieve_flipped": info["flipped"], "retrieve_negativity": info["negativity"],
           "round_trip_residual": resid}
    log(f"two-photon: kernel {1e3 * kern:.1f} us, call {1e3 * call:.2f} ms; retrieveSolution {1e3 * t_ret:.1f} ms "
        f"(flipped {info['flipped']}, round-trip residual {resid:.3g})")
    return out


def fit_leg(tp, truth, num):
    """fitStructures_full (data_processing.jl:26-46) on the device: the 38-shell
    structure against a rotated copy of itself, L = 18, LMAX = 25, similarity
    over shells 1..38 for every one of (num+1)^3 Euler triples of the
    reference's default ranges; the rotation and similarity kernels timed with
    HIP events, the call end to end (quaternions, upload, both kernels, copy)."""
    from tp_amd import fit
    L, K = 18, truth.KMAX
    b = tp.HIP_calculate_basis(L, truth.LMAX, 2, 1)
    t, p, g = np.pi / 4, np.pi / 2, 7 * np.pi / 20      # a point of the num = 20 grid
    vol = fit.rotateStructure(truth, -t, -g, -p, K, L, b)
    ctx = tp.HIP_init()
    sf = fit.StructureFit(b, truth, K).set_volume(vol)
    sf.similarity(np.zeros((64, 3)))
    th = np.linspace(0.0, np.pi / 2, num + 1)
    ph = np.linspace(0.0, np.pi, num + 1)
    T, P, G = np.meshgrid(th, ph, ph, indexing="ij")
    euler = np.stack([T.ravel(), P.ravel(), G.ravel()], axis=1)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    sc = sf.similarity(euler)
    wall = time.perf_counter() - t0
    rot_ms, sim_ms = ctx.last_timings()[:2]
    ctx.set_timing(False)
    i = int(np.argmax(sc))
    n = len(euler)
    B = truth.LMAX
    flops = 8.0 * 2 * B * (B * (B + 1) / 2) * K * n      # 4 FMAs per (theta_j, |m|, l) per shell
    out = {"workload": f"fitStructures_full, num={num}: {n} rotations x {K} shells, L={L}, LMAX={B}",
           "rotations": n, "call_s": wall, "rotations_per_s": n / wall, "rotate_kernel_ms": rot_ms,
           "similarity_kernel_ms": sim_ms, "similarity_tflops": flops / (sim_ms * 1e-3) / 1e12,
           "best": float(sc[i]), "best_euler": [float(x) for x in euler[i]]}
    log(f"fit: {n} rotations in {1e3 * wall:.1f} ms ({n / wall:.0f}/s; rotate {rot_ms:.2f} ms, similarity "
        f"{sim_ms:.2f} ms = {out['similarity_tflops']:.1f} TFLOP/s), best {sc[i]:.12f} at {np.round(euler[i], 4)}")
    return out


def generate_leg(tp, truth, npic):
    """generateHistogram (datagen.jl:363-415) on the device: npic pictures of
    the 38-shell intensity (pointsPerOrientation: 1000 incident photons,
    envelope qcut/3, qcut = the volume's qmax, lambda = 0, as in
    jobs/run_generate_histograms.jl) histogrammed at N = 32, K2 = 38, K3 = 26
    without leaving the device; picture synthesis and histogram kernels timed
    with HIP events."""
    from tp_amd import pictures
    ctx = tp.HIP_init()
    qcut = truth.rmax
    h = tp.PhotonHistogram(32, 38, 26, qcut / 38)
    pictures.generate_histograms(h, truth, qcut, 256, photons_per_image=1000, lam=0.0, capacity=512, seed=3)
    h = tp.PhotonHistogram(32, 38, 26, qcut / 38)
    ctx.set_timing(True)
    t0 = time.perf_counter()
    over = pictures.generate_histograms(h, truth, qcut, npic, photons_per_image=1000, lam=0.0, capacity=512, seed=7)
    wall = time.perf_counter() - t0
    gen_ms, hist_ms = ctx.last_timings()[:2]
    ctx.set_timing(False)
    r = h.read()
    n1, n3 = int(r["n1"].sum()), int(r["n3"].sum())
    out = {"workload": f"{npic} pictures of the {truth.KMAX}-shell intensity (LMAX={truth.LMAX}), 1000 incident "
                       f"photons, N=32, K2=38, K3=26, lambda=0",
           "call_s": wall, "pictures_per_s": npic / wall, "synthesis_ms": gen_ms, "histogram_ms": hist_ms,
           "photons": n1, "photons_per_picture": n1 / npic, "triplets": n3, "overflow": over}
    log(f"generateHistogram: {npic} pictures in {wall:.3f} s ({npic / wall:.3g}/s; synthesis {gen_ms:.1f} ms, "
        f"histograms {hist_ms:.1f} ms), {n1 / npic:.1f} photons/picture, {n3} triplets")
    return out


def ewald_leg(cfg, tp, start, c3ref, reps, workload=None, truth=None, batch_sweeps=0):
    """The plain FP64 evaluation on the curved Ewald sphere: the full-size
    problem (L = 18, K = 26, N = 32, the same start structure and histogram)
    with a basis at lambda = 2.5 (the reference's own GPU test value) and
    dq = pi / 38; synchronous c3 + energy evaluations from the host, stage
    times with HIP events.  With batch_sweeps > 0 also the batched form on
    the same basis (FP16X3_T64): the 20 runs of the batched leg (K_r from
    seeds 1000..1019, 64 proposals each) scored against the full histogram
    cut to K_r shells, sweeps timed after one untimed sweep."""
    ctx = tp.HIP_init()
    lam, dq = 2.5, np.pi / 38
    b = tp.HIP_calculate_basis(cfg["L"], cfg["LMAX"], cfg["N"], cfg["K"], lam, dq)
    c3ref_full = c3ref
    c3ref = b.histogram(c3ref)   # on the device before the timed region
    tp.energy(start, b, c3ref)
    t0 = time.perf_counter()
    for _ in range(reps):
        E = tp.energy(start, b, c3ref)
    el = time.perf_counter() - t0
    ctx.set_timing(True)
    tp.energy(start, b, c3ref)
    st = ctx.last_timings()
    ctx.set_timing(False)
    out = {"workload": f"L={cfg['L']}, K={cfg['K']}, N={cfg['N']}, lambda={lam}, dq=pi/38", "evals_per_s": reps / el,
           "ms_per_eval": 1e3 * el / reps, "energy": E,
           "stage_ms": {"triple": st[0], "glam_G": st[1], "c3_energy": st[2], "final": st[3]}}
    log(f"curved Ewald sphere: {reps / el:.1f} evals/s ({1e3 * el / reps:.3f} ms), stage ms {np.round(st, 4)}")
    # the same evaluation in FP16X3_T64 (mixed-precision triple products)
    b.set_precision("fp16x3_t64")
    E16 = tp.energy(start, b, c3ref)
    t0 = time.perf_counter()
    for _ in range(reps):
        tp.energy(start, b, c3ref)
    el16 = time.perf_counter() - t0
    out["fp16x3_t64"] = {"evals_per_s": reps / el16, "energy_abs_err_vs_fp64": abs(E16 - E)}
    if batch_sweeps > 0:
        seeds = tuple(range(1000, 1020))
        Ks = [int(np.random.default_rng(s).integers(10, 27)) for s in seeds]
        starts = [workload.random_start_structure(truth, cfg["L"], s) for s in seeds]
        refs = [np.ascontiguousarray(c3ref_full[:, :, :K, :K, :K]) for K in Ks]
        nprop = 64
        B = tp.AnnealBatch(b, refs, starts, seeds, nprop=nprop, hist_cap=batch_sweeps + 8)
        B.probe(0.1)
        E0 = [B.group(g)["E"] for g in range(len(Ks))]
        B.run(1, factor=SWEEP_FACTOR)   # graph capture, untimed
        t0 = time.perf_counter()
        B.run(batch_sweeps, factor=SWEEP_FACTOR)
        elb = time.perf_counter() - t0
        ne = len(Ks) * nprop
        out["batched_fp16x3_t64"] = {
            "runs": len(Ks), "K": Ks, "proposals_per_run": nprop, "evals_per_sweep": ne, "sweeps": batch_sweeps,
            "ms_per_sweep": 1e3 * elb / batch_sweeps, "evals_per_s": ne * batch_sweeps / elb,
            "E0": E0, "E_final": [B.group(g)["E"] for g in range(len(Ks))]}
        log(f"curved Ewald sphere, batched: {1e3 * elb / batch_sweeps:.2f} ms/sweep, "
            f"{ne * batch_sweeps / elb:.0f} evals/s")
        del B
    b.set_precision("fp64")
    return out


def structure_leg(tp, reps):
    """createSphericalHarmonicsStructure (structure.jl:124-134) of crambin
    (the ATOM records the reference's structure test loads, 670 atoms) at
    LMAX = 35, KMAX = 38, rmax = 38, as the histogram-generation runs make
    their intensity: density sampling, SH analysis, spherical Bessel
    transform, squared modulus and analysis again, all on the device (HIP
    events), and the density-sampling kernel alone.  Its pair rate counts all
    atom-sample pairs, including those the radial cull never evaluates."""
    from tp_amd import structure
    ctx = tp.HIP_init()
    LMAX, K, rmax = 35, 38, 38.0
    atoms = structure.atom_table(structure.loadPDB(STRUCTURE_PDB))
    structure.createSphericalHarmonicsStructure(atoms, LMAX, K, rmax)
    ctx.set_timing(True)
    dev, wall = [], []
    for _ in range(reps):
        t0 = time.perf_counter()
        structure.createSphericalHarmonicsStructure(atoms, LMAX, K, rmax)
        wall.append(time.perf_counter() - t0)
        dev.append(ctx.last_timings()[0])
    dens = []
    for _ in range(reps):
        structure.getDensitySurfaces(atoms, LMAX, K, rmax)
        dens.append(ctx.last_timings()[0])
    ctx.set_timing(False)
    d_ms = float(np.median(dens))
    pairs = len(atoms) * K * (2 * LMAX) ** 2
    out = {"workload": f"crambin ({len(atoms)} atoms), LMAX={LMAX}, KMAX={K}, rmax={rmax}",
           "call_ms": 1e3 * float(np.median(wall)), "device_ms": float(np.median(dev)),
           "density_kernel_ms": d_ms, "atom_sample_pairs_covered_per_s": pairs / (d_ms * 1e-3)}
    log(f"createSphericalHarmonicsStructure: call {out['call_ms']:.2f} ms, device {out['device_ms']:.3f} ms "
        f"(density sampling {d_ms:.3f} ms, all {pairs:.3g} atom-sample pairs covered at "
        f"{out['atom_sample_pairs_covered_per_s']:.3g}/s, radially culled)")
    return out


def phase_leg(tp, truth, iters):
    """sRAAR (phases.jl:31-95) of the 38-shell intensity (LMAX = 25, 50 x 50
    surface grid per shell) with the reference's post-processing settings
    (data_processing.jl:508: beta 0.75 -> 0.9, tau = 350, cutoff_factor 0.3),
    every transform, projection and reflection on the device; initial phases
    drawn from a fixed seed."""
    from tp_amd import phases
    phases.sRAAR(truth, 2, 0.75, 0.9, 350.0, 0.3, seed=1)
    t0 = time.perf_counter()
    vol, diff = phases.sRAAR(truth, iters, 0.75, 0.9, 350.0, 0.3, seed=1, return_diff=True)
    wall = time.perf_counter() - t0
    out = {"workload": f"sRAAR of the {truth.KMAX}-shell intensity, LMAX={truth.LMAX}, {iters} iterations "
                       f"(beta 0.75 -> 0.9, tau 350, cutoff 0.3)",
           "seconds": wall, "ms_per_iteration": 1e3 * wall / max(iters, 1),
           "diff_first": float(diff[0]) if iters else None, "diff_last": float(diff[-1]) if iters else None}
    log(f"sRAAR: {iters} iterations in {wall:.3f} s ({out['ms_per_iteration']:.3f} ms each); "
        f"change {out['diff_first']:.3g} -> {out['diff_last']:.3g}")
    return out


def gradient_leg(tp, basis, c3ref, start, plain, nsteps, ncand=16):
    """Gradient-guided moves (tp_amd.descent) on the plain run's problem (K=26,
    its start structure and histogram): per step one FP64 energy + gradient
    (tp_energy_grad), a ladder of ncand step lengths along the Riemannian
    steepest descent on every SO(2l+1) scored in one batched FP16 hi/lo
    evaluation, the best accepted on its FP64 energy.  The candidate batch's
    plan and graph are built before the timed region (as the batched runs'
    graphs are)."""
    from tp_amd import descent
    basis.set_precision("fp16x3")
    tp.energy_batch([start.coeff] * ncand, [basis] * ncand, [c3ref] * ncand)
    tp.energy_grad(start.coeff, basis, c3ref)
    E_plain, t_plain = plain["E_final"], plain["seconds"]
    r = descent.gradient_anneal(start, basis, c3ref, nsteps, ncand=ncand)
    h = r["history"]
    ts, Es = np.array([t for t, _ in h]), np.array([e for _, e in h])
    per = ts[-1] / max(1, len(h) - 1)
    reach = _time_to(Es, lambda i: ts[i], E_plain)
    within = _time_to(Es, lambda i: ts[i], E_plain * 1.001)
    at_plain = Es[np.searchsorted(ts, t_plain, side="right") - 1]
    out = {"mode": "FP64 energy + gradient (tp_energy_grad) per step; %d step lengths along exp(-eta Omega_l) "
                   "scored in one batched FP16x3 evaluation; best accepted on its FP64 energy (T = 0)" % ncand,
           "steps": len(h) - 1, "ms_per_step": 1e3 * per, "E0": float(Es[0]), "E_final": float(Es[-1]),
           "E_trace": [float(e) for e in Es], "step_lengths": r["step_lengths"],
           "t_reach_plain_final": reach, "t_within_0.1pct": within,
           "E_at_plain_seconds": float(at_plain), "plain_seconds": t_plain, "plain_E_final": E_plain,
           "speedup_reach_plain_final": (t_plain / reach) if reach else None}
    log(f"gradient moves: {len(h) - 1} steps, {1e3 * per:.2f} ms/step, E {Es[0]:.4f} -> {Es[-1]:.4f}; "
        f"reaches the plain run's final {E_plain:.4f} at {1e3 * reach if reach else float('nan'):.2f} ms "
        f"(plain: {1e3 * t_plain:.1f} ms)")
    basis.set_precision("fp64")
    return out


def learned_leg(cfg, tp, workload, truth, basis, c3ref, H, start, plain, samples, train_steps, cmp_steps,
                tte_steps, nprop=64, nlearn=48, seeds=tuple(range(1000, 1020)), prec="fp16x3"):
    """Learned proposer (tp_amd.descent.learned_gradient_runs), 48 of 64 slots
    per run, the reference's random move in the other 16:
    (1) move log from runs shaped as the concurrent-runs workload (20 runs,
        K_r ~ U{10..26} from seeds 1000..1019, 64 candidates per step): greedy
        gradient steps with random per-l profiles, in chunks of 50 steps from
        fresh start structures (randomStartStructure(seed + 20 j)), until the
        log holds `samples` pairs;
    (2) device training of the MLP on the log;
    (3) the 20 runs from randomStartStructure(seed), cmp_steps steps each of
        the batched random-walk sweep (64 plane moves, graph replays; the
        method without the learned part), the learned proposer, and the
        learned proposer with random profiles in place of the network's;
    (4) time-to-energy on the plain run's problem (K=26, its start structure
        and histogram) against the plain FP64 anneal.
    Batches (in mixed precision prec) and graphs are built before each
    timed region."""
    from tp_amd import descent
    Ks = [int(np.random.default_rng(s).integers(10, 27)) for s in seeds]
    refs = [basis.histogram(np.ascontiguousarray(c3ref[:, :, :K, :K, :K]), range(1, K + 1)) for K in Ks]
    nrand = nprop - nlearn
    out = {"nprop": nprop, "nlearn": nlearn, "nrand": nrand, "K": Ks, "batch_precision": prec}
    # (1) data
    mlog = descent.MoveLog(samples)
    steps, j, t0 = 0, 0, time.perf_counter()
    while not mlog.full() and j < 40:
        starts = [workload.random_start_structure(truth, cfg["L"], s + 20 * j) for s in seeds]
        batch = _geodesic_batch(tp, basis, refs, starts, nprop, prec)
        _, n = descent.learned_gradient_runs(starts, basis, refs, 50, nlearn=nlearn, nrand=nrand, seed=j, log=mlog,
                                             batch=batch)
        steps += n
        j += 1
        log(f"learned: data chunk {j}: {steps} steps, {mlog.n} samples, {time.perf_counter() - t0:.1f}s")
    t_data = time.perf_counter() - t0
    out["data"] = {"samples": mlog.n, "steps": steps, "runs": len(seeds), "chunks": j, "seconds": t_data,
                   "samples_are": "(shell features, per-l step profile) of the 8 best geodesic candidates per run "
                                  "and step that pass the Metropolis test, one pair per shell of the run"}
    # (2) training
    X, T = mlog.arrays()
    mlp = tp.MlpProposer(seed=1)
    t0 = time.perf_counter()
    tr = mlp.train(X, T, train_steps, batch=512, lr=1e-3, seed=2)
    t_train = time.perf_counter() - t0
    prof = np.log(descent.ratios_of(mlp.forward(X[:4096]), len(range(2, cfg["L"] + 1, 2))))
    out["train"] = {"steps": train_steps, "batch": 512, "seconds": t_train, "samples_seen": train_steps * 512,
                    "loss_first": float(tr[:20].mean()), "loss_last": float(tr[-20:].mean()),
                    "predicted_log_profile_mean": [float(x) for x in prof.mean(0)],
                    "predicted_log_profile_std": [float(x) for x in prof.std(0)]}
    log(f"learned: trained {train_steps} steps in {t_train:.2f}s, loss {tr[:20].mean():.4f} -> {tr[-20:].mean():.4f}")
    del mlog, X, T
    # (3) the concurrent runs with and without the learned part
    starts = [workload.random_start_structure(truth, cfg["L"], s) for s in seeds]
    cmp = {}
    basis.set_precision(prec)
    B = tp.AnnealBatch(basis, refs, starts, list(seeds), nprop=nprop, hist_cap=cmp_steps + 8)
    B.probe(0.1)
    B.run(0, factor=SWEEP_FACTOR)
    E0 = np.array([B.group(g)["E"] for g in range(len(seeds))])
    t0 = time.perf_counter()
    B.run(cmp_steps, factor=SWEEP_FACTOR)
    el = time.perf_counter() - t0
    dE = np.array([B.group(g)["E"] for g in range(len(seeds))]) - E0
    cmp["random_walk_batched"] = {"ms_per_step": 1e3 * el / cmp_steps, "mean_dE": float(dE.mean()),
                                  "dE": [float(x) for x in dE],
                                  "mode": "64 plane moves per run (|b| <= pi/4 Haar frame), best-of-64 + Metropolis "
                                          "at the probed temperature, one graph replay per step"}
    del B
    basis.set_precision("fp64")
    for name, m in (("learned", mlp), ("random_profile", None)):
        batch = _geodesic_batch(tp, basis, refs, starts, nprop, prec)
        t0 = time.perf_counter()
        res, n = descent.learned_gradient_runs(starts, basis, refs, cmp_steps, mlp=m, nlearn=nlearn, nrand=nrand,
                                               seed=77, batch=batch)
        el = time.perf_counter() - t0
        dE = np.array([r["E"] - r["history"][0][1] for r in res])
        cmp[name] = {"ms_per_step": 1e3 * el / n, "mean_dE": float(dE.mean()), "dE": [float(x) for x in dE],
                     "geodesic_wins": int(sum(r["wins"]["geodesic"] for r in res)),
                     "plane_wins": int(sum(r["wins"]["plane"] for r in res))}
    cmp["steps"] = cmp_steps
    cmp["gain_vs_random_walk"] = cmp["learned"]["mean_dE"] / cmp["random_walk_batched"]["mean_dE"]
    cmp["gain_vs_random_profile"] = cmp["learned"]["mean_dE"] / cmp["random_profile"]["mean_dE"]
    out["compare"] = cmp
    log(f"learned: mean dE over {cmp_steps} steps: random walk {cmp['random_walk_batched']['mean_dE']:.3f}, "
        f"learned {cmp['learned']['mean_dE']:.3f}, random profile {cmp['random_profile']['mean_dE']:.3f}")
    # (4) time-to-energy on the plain run's problem
    E_plain, t_plain = plain["E_final"], plain["seconds"]
    tte = {}
    for name, m in (("learned", mlp), ("random_profile", None)):
        batch = _geodesic_batch(tp, basis, [H], [start], nprop, prec)
        r = descent.learned_gradient_anneal(start, basis, H, tte_steps, mlp=m, nlearn=nlearn, nrand=nrand, seed=1,
                                            batch=batch)
        h = r["history"]
        ts, Es = np.array([t for t, _ in h]), np.array([e for _, e in h])
        tte[name] = {"steps": len(h) - 1, "ms_per_step": 1e3 * ts[-1] / max(1, len(h) - 1), "E_final": float(Es[-1]),
                     "E_trace": [float(e) for e in Es],
                     "t_within_0.1pct": _time_to(Es, lambda i: ts[i], E_plain * 1.001),
                     "t_reach_plain_final": _time_to(Es, lambda i: ts[i], E_plain)}
    Ep = np.asarray(plain["E_trace"])
    per_step = t_plain / len(Ep)
    tte["plain_fp64"] = {"steps": len(Ep), "seconds": t_plain, "E_final": E_plain,
                         "t_within_0.1pct": _time_to(Ep, lambda i: (i + 1) * per_step, E_plain * 1.001),
                         "t_reach_plain_final": t_plain}
    lf = tte["learned"]["t_reach_plain_final"]
    tte["speedup_reach_plain_final"] = (t_plain / lf) if lf else None
    tte["E0"] = float(plain["E0"])
    out["time_to_energy"] = tte
    log(f"learned: reaches the plain run's final {E_plain:.4f} at "
        f"{1e3 * lf if lf else float('nan'):.2f} ms (plain {1e3 * t_plain:.1f} ms; speedup "
        f"{tte['speedup_reach_plain_final']}); E after {tte_steps} steps {tte['learned']['E_final']:.4f}")
    return out


def _geodesic_batch(tp, basis, refs, starts, nprop, prec):
    """A batch of nprop slots per run for the learned proposer, made in mixed precision prec."""
    old = basis.precision()
    basis.set_precision(prec)
    try:
        return tp.AnnealBatch(basis, list(refs), list(starts), list(range(len(refs))), nprop=nprop)
    finally:
        basis.set_precision(old)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--anneal-steps", type=int, default=2000)
    ap.add_argument("--cpu-cols", type=int, default=1600)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--batch-sweeps", type=int, default=20, help="sweeps of the concurrent-runs leg (0: skip)")
    ap.add_argument("--learned-samples", type=int, default=1_000_000,
                    help="logged training pairs of the learned proposer (0: skip the learned leg)")
    ap.add_argument("--learned-data-sweeps", type=int, default=None,
                    help="0 skips the learned leg (the older form of --learned-samples 0)")
    ap.add_argument("--learned-precision", default=None, choices=("fp16x3", "fp16x3_t64"),
                    help="batch mode of the learned leg (default: the faster mode of the batched leg, else fp16x3)")
    ap.add_argument("--learned-train-steps", type=int, default=4000)
    ap.add_argument("--learned-cmp-steps", type=int, default=20)
    ap.add_argument("--learned-tte-steps", type=int, default=40)
    ap.add_argument("--grad-steps", type=int, default=30, help="steps of the gradient-guided moves (0: skip)")
    ap.add_argument("--hist-pictures", type=int, default=32768, help="pictures of the histogram leg (0: skip)")
    ap.add_argument("--c2-reps", type=int, default=20, help="timed twoPhotons calls of the two-photon leg (0: skip)")
    ap.add_argument("--fit-num", type=int, default=20, help="grid divisions of the rotational-fit leg (0: skip)")
    ap.add_argument("--sraar-iters", type=int, default=1001, help="iterations of the phase-retrieval leg (0: skip)")
    ap.add_argument("--gen-pictures", type=int, default=65536, help="pictures of the histogram-generation leg (0: skip)")
    ap.add_argument("--structure-reps", type=int, default=5, help="timed calls of the structure leg (0: skip)")
    ap.add_argument("--ewald-reps", type=int, default=50, help="timed evaluations at lambda = 2.5 (0: skip)")
    ap.add_argument("--ewald-batch-sweeps", type=int, default=5,
                    help="timed sweeps of the batched runs at lambda = 2.5 (0: skip)")
    ap.add_argument("--grad-profile", type=int, default=0,
                    help="instead of the benchmark: time the parts of a gradient move, then this many gradient calls")
    ap.add_argument("--precision-study", type=int, default=0,
                    help="instead of the benchmark: c3/energy errors of the mixed-precision modes against FP64 "
                         "for this many random orientations of the full-size structure")
    args = ap.parse_args()
    if args.precision_study:
        precision_study(args.precision_study)
        return
    if args.grad_profile:
        gradient_profile(args.grad_profile)
        return
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: there is no multi-GPU path (sharding over k-triple "
              "columns is designed, not built); run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    import tp_amd as tp
    from tp_amd import workload
    cfg = workload.CONFIGS["full"]
    ctx = tp.HIP_init()
    truth, basis, c3ref, start = setup(cfg, tp, workload)
    H = basis.histogram(c3ref)   # the histogram on the device, uploaded outside the timed region
    state = tp.DeviceState(start, basis)
    state.energy(H)
    # GPU clock warm-up, untimed: the host-side set-up above leaves the GPU idle
    # for about a second; 0.3 s of the same evaluations brings it back to its
    # working clock before the W warm-up steps and the K timed ones
    t_end = time.perf_counter() + 0.3
    while time.perf_counter() < t_end:
        state.energy(H)

    for _ in range(args.warmup):
        state.energy(H)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        E = state.energy(H)
    elapsed = time.perf_counter() - t0
    evals_per_s = args.steps / elapsed
    kern_ms = stage_times(ctx, state, H, args.steps)
    n_m = folded_classes(cfg["L"])
    flops_g1 = 2.0 * n_m * basis.n_lt * basis.klength         # G = H . As
    # c3 = Bm . G and the energy on the folded separated (alpha, beta) grid
    flops_g2 = sep_grid_flops(cfg["L"], cfg["N"]) * basis.klength
    K = cfg["K"]
    # triple products on sorted l-triples (the form k_triple_sym executes); the
    # coupling form over all 505 ordered l-triples is reported beside it
    flops_tri = sorted_lt_flops(cfg["L"], K)
    flops_coupling = 8.0 * (K * (K + 1) // 2 * basis.basislen + basis.klength * lt_n3_sum(cfg["L"]))
    achieved = flops_tri / (kern_ms[0] * 1e-3) / 1e12
    log(f"evals/s {evals_per_s:.1f}  stage ms: triple {kern_ms[0]:.4f} G {kern_ms[1]:.4f} c3+energy "
        f"{kern_ms[2]:.4f} final {kern_ms[3]:.4f}  E={E:.10g}")

    # plain annealing (2000 steps), device resident
    astate = tp.DeviceState(start, basis)
    t0 = time.perf_counter()
    res = tp.rotate_all_at_once(astate, basis, H, args.anneal_steps, np.pi / 4, 1.01, 0.1, 0.99998, seed=1000,
                                ntemp=50)
    t_anneal = time.perf_counter() - t0
    steps_per_s = (args.anneal_steps + 50 + 1) / t_anneal
    log(f"anneal: {args.anneal_steps} steps in {t_anneal:.2f}s, E0 {res['E0']:.10g} -> {res['E_final']:.10g}, "
        f"accepted {int(res['accepted'].sum())}")

    # mixed precision (FP16 hi/lo compensated contraction), same structure and histogram
    basis.set_precision("fp16x3")
    E16 = state.energy(H)
    for _ in range(args.warmup):
        state.energy(H)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        state.energy(H)
    el16 = time.perf_counter() - t0
    kern16 = stage_times(ctx, state, H, args.steps)
    basis.set_precision("triple16")
    Et16 = state.energy(H)
    for _ in range(args.warmup):
        state.energy(H)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        state.energy(H)
    elt16 = time.perf_counter() - t0
    kernt16 = stage_times(ctx, state, H, args.steps)
    basis.set_precision("fp16x3_t64")
    Et64 = state.energy(H)
    for _ in range(args.warmup):
        state.energy(H)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        state.energy(H)
    elt64 = time.perf_counter() - t0
    kernt64 = stage_times(ctx, state, H, args.steps)
    basis.set_precision("fp16x3")
    mstate = tp.DeviceState(start, basis)
    t0 = time.perf_counter()
    res16 = tp.rotate_all_at_once(mstate, basis, H, args.anneal_steps, np.pi / 4, 1.01, 0.1, 0.99998, seed=1000,
                                  ntemp=50)
    t_anneal16 = time.perf_counter() - t0
    basis.set_precision("fp64")
    log(f"fp16x3: evals/s {args.steps / el16:.1f}  stage ms {np.round(kern16, 4)}  |dE| {abs(E16 - E):.3g}  "
        f"anneal {t_anneal16:.2f}s -> {res16['E_final']:.10g}")

    batch = batch_t64 = None
    batch_prec = "fp16x3"
    if args.batch_sweeps > 0:
        batch = batch_leg(cfg, tp, workload, truth, basis, c3ref, args.batch_sweeps)
        batch_t64 = batch_leg(cfg, tp, workload, truth, basis, c3ref, args.batch_sweeps, prec="fp16x3_t64")
        if batch_t64["ms_per_sweep"] < batch["ms_per_sweep"]:
            batch_prec = "fp16x3_t64"
    hist = hist_leg(tp, workload, args.hist_pictures) if args.hist_pictures > 0 else None
    c2 = c2_leg(tp, truth, basis, args.c2_reps) if args.c2_reps > 0 else None
    fitr = fit_leg(tp, truth, args.fit_num) if args.fit_num > 0 else None
    phase = phase_leg(tp, truth, args.sraar_iters) if args.sraar_iters > 0 else None
    gen = generate_leg(tp, truth, args.gen_pictures) if args.gen_pictures > 0 else None
    struct = structure_leg(tp, args.structure_reps) if args.structure_reps > 0 else None
    ewald = (ewald_leg(cfg, tp, start, c3ref, args.ewald_reps, workload, truth, args.ewald_batch_sweeps)
             if args.ewald_reps > 0 else None)
    plain = {"E_final": res["E_final"], "E0": res["E0"], "seconds": t_anneal, "E_trace": res["E"]}
    grad = None
    if args.grad_steps > 0:
        grad = gradient_leg(tp, basis, H, start, plain, args.grad_steps)
    learned = None
    if args.learned_data_sweeps == 0:
        args.learned_samples = 0
    if args.learned_precision:
        batch_prec = args.learned_precision
    if args.learned_samples > 0:
        learned = learned_leg(cfg, tp, workload, truth, basis, c3ref, H, start, plain, args.learned_samples,
                              args.learned_train_steps, args.learned_cmp_steps, args.learned_tte_steps,
                              prec=batch_prec)

    out = {
        "metric": METRIC,
        "value": evals_per_s,
        "unit": "c3+energy evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * elapsed / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {
            "workload": "plain FP64 c3 + Bayes energy: 1000 random point atoms (r=26), KMAX=38, L=18, K=26, "
                        "N=32 -> 3276 k-triples x 2048 bins, 505 l-triples; 1e9-count Poisson histogram",
            "form": "triple products on sorted l-triples (3j column symmetry); G = H.As and c3 = Bm.G on FP64 MFMA "
                    "with the energy fused",
        },
        "mc_steps_per_s": steps_per_s,
        "mc_anneal": {"steps": args.anneal_steps, "temperature_probes": 50, "seconds": t_anneal,
                      "E0": res["E0"], "E_final": res["E_final"], "accepted": int(res["accepted"].sum())},
        "kernel_ms": {"triple": kern_ms[0], "gemm_G": kern_ms[1], "c3sep_energy": kern_ms[2],
                      "final": kern_ms[3]},
        "mixed_precision": {
            "mode": "fp16x3: sorted l-triple products (FP32 stage 1, FP16 hi/lo stage 2), c3 = F.As on FP16 MFMA, "
                    "hi.hi + (hi.lo + lo.hi), FP32 accumulation, FP64 energy",
            "evals_per_s": args.steps / el16, "ms_per_eval": 1e3 * el16 / args.steps,
            "stage_ms": {"triple": kern16[0], "split": kern16[1], "gemm16_c3_energy": kern16[2], "final": kern16[3]},
            "gemm16_tflops_algorithmic": 2.0 * basis.R * basis.n_lt * basis.klength / (kern16[2] * 1e-3) / 1e12,
            "energy_abs_err_vs_fp64": abs(E16 - E), "mc_steps_per_s": (args.anneal_steps + 51) / t_anneal16,
            "mc_E_final": res16["E_final"],
            "triple16": {"mode": "compensated FP16 triple products, FP64 separated-grid contraction",
                         "evals_per_s": args.steps / elt16, "energy_abs_err_vs_fp64": abs(Et16 - E),
                         "stage_ms": {"triple": kernt16[0], "gemm_G": kernt16[1], "c3_energy": kernt16[2],
                                      "final": kernt16[3]}},
            "fp16x3_t64": {"mode": "triple products on the sorted l-triples (FP32 MFMA stage 1, FP16 hi/lo MFMA "
                                   "stage 2), G = H.As on FP16 hi/lo MFMA (k_hgemm_g), c3 = Bm.G and the energy "
                                   "on the FP64 separated grid",
                           "evals_per_s": args.steps / elt64, "energy_abs_err_vs_fp64": abs(Et64 - E),
                           "stage_ms": {"triple": kernt64[0], "gemm_G_fp16x3": kernt64[1],
                                        "c3sep_energy": kernt64[2], "final": kernt64[3]}}},
        "roofline": {"bound": "mfma", "kernel": "k_triple_sym (triple products on the 125 sorted l-triples, "
                                               "FP64 MFMA)",
                     "achieved": achieved, "peak": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP64_PEAK_TFLOPS,
                     "traffic": TRIPLE_FETCH_KIB * 1024 * FETCH_SCALE + TRIPLE_WRITE_KIB * 1024 * WRITE_SCALE,
                     "flops_per_launch": flops_tri, "launch_ms": kern_ms[0],
                     "coupling_form_flops": flops_coupling,
                     "coupling_form_tflops": flops_coupling / (kern_ms[0] * 1e-3) / 1e12,
                     "c3sep_energy_tflops": flops_g2 / (kern_ms[2] * 1e-3) / 1e12,
                     "gemm_G_tflops": flops_g1 / (kern_ms[1] * 1e-3) / 1e12},
    }
    if batch is not None:
        out["batched"] = batch
        out["batched_fp16x3_t64"] = batch_t64
    if learned is not None:
        out["learned"] = learned
    if grad is not None:
        out["gradient_moves"] = grad
    if hist is not None:
        out["histograms"] = hist
    if c2 is not None:
        out["two_photon"] = c2
    if fitr is not None:
        out["rotational_fit"] = fitr
    if phase is not None:
        out["phase_retrieval"] = phase
    if gen is not None:
        out["generated_histograms"] = gen
    if struct is not None:
        out["structure"] = struct
    if ewald is not None:
        out["curved_ewald"] = ewald
    if not args.no_cpu:
        # every core this process may run on, but no more threads than the
        # OpenMP setting of the machine (a GPU box grants its share of the
        # host's cores and sets OMP_NUM_THREADS to that share)
        affinity = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
        omp = int(os.environ.get("OMP_NUM_THREADS", "0") or 0)
        nthreads = min(affinity, omp) if omp > 0 else affinity
        cb = cpu_baseline(cfg, start, c3ref, args.cpu_cols, nthreads)
        cb["host"] = {"nproc": os.cpu_count(), "affinity_cores": affinity, "OMP_NUM_THREADS": omp or None,
                      "threads_used": nthreads, "model": cpu_model()}
        out["cpu_baseline"] = cb
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
