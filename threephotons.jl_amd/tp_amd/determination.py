"""Structure determination driver over the device path (determination.jl).

  loadHistograms(K2, K3, file, load_c1)     datagen.jl:432-475 (c3 part: tp_load_histograms)
  write_histograms / read_histograms        the histogram file (the reference serialises
                                            (params, c2, c3, c1) with Julia's serialize,
                                            datagen.jl:424; here a numpy .npz, see below)
  randomStartStructure(volume, K, L, seed)  determination.jl:4-17
  AnnealRun                                 rotate_all_at_once's `state` for one L
                                            (determination.jl:281-338) on tp_anneal_begin /
                                            tp_anneal_steps, with saveState's checkpoint
                                            (:264-278) as tp_checkpoint_save / _load
  rotate_all_at_once(...)                   the annealing loop until the step size of the
                                            largest l is at or below 1.5 degrees (:1, :322),
                                            a checkpoint every 500 steps (regular_action)
  rotation_search(params, workdir)          determination.jl:154-229: histograms -> start
                                            structure (retrieveSolution, randomStartStructure)
                                            -> optimizer, resumable from its checkpoint

Histogram file: a .npz (numpy.savez, no pickled objects) with arrays "c1" (K2),
"c2" (N, K2, K2) and "c3" (N, 2N or N, K3, K3, K3), all float64 and indexed as
the reference's arrays ([a, k2, k1], [a, b, k3, k2, k1], zero-based), plus
"params": the reference's params dict as a JSON string.
"""
import ctypes as C
import json
import math
import os

import numpy as np

from . import (DeviceState, HIP_calculate_basis, SphericalHarmonicsVolume, check, lib, retrieveSolution)
from ._lib import AnnealParams, AnnealStatus, CheckpointInfo

# determination.jl:1: 1.5 degrees, the step size of the largest l at which the run stops
minstepsize = 1.5 * math.pi / 180.0
CHECKPOINT_EVERY = 500   # regular_action (determination.jl:264-278)


# ---- histograms -------------------------------------------------------------

def write_histograms(path, params, c2, c3, c1):
    """Write (params, c2, c3, c1) as generateHistogram's file holds them (datagen.jl:424)."""
    c1 = np.asarray(c1, dtype=np.float64)
    c2 = np.asarray(c2, dtype=np.float64)
    c3 = np.asarray(c3, dtype=np.float64)
    if c2.ndim != 3 or c3.ndim != 5 or c3.shape[0] != c2.shape[0]:
        raise ValueError("c2 must be (N, K2, K2) and c3 (N, 2N or N, K3, K3, K3)")
    with open(path, "wb") as f:
        np.savez(f, params=np.array(json.dumps(params)), c1=c1, c2=c2, c3=c3)


def read_histograms(path):
    """(params, c2, c3, c1) from a file of write_histograms."""
    with np.load(path, allow_pickle=False) as z:
        return json.loads(str(z["params"])), z["c2"], z["c3"], z["c1"]


def load_c3(c3_full, K3):
    """loadHistograms' treatment of the triplet histogram (tp_load_histograms):
    doubling of an old N x N histogram, clamp to >= 1e-30, the first K3 shells,
    divided by their total.  Returns (N, 2N, K3, K3, K3)."""
    c3_full = np.asarray(c3_full, dtype=np.float64)
    if c3_full.ndim != 5:
        raise ValueError("c3 must be (N, N or 2N, Kf, Kf, Kf)")
    NA, NB, Kf = c3_full.shape[0], c3_full.shape[1], c3_full.shape[2]
    flat = np.ascontiguousarray(c3_full.ravel(order="F"))
    out = np.zeros(2 * NA * NA * K3 ** 3)
    check(lib().tp_load_histograms(flat.ctypes.data, NA, NB, Kf, K3, out.ctypes.data), "tp_load_histograms")
    return out.reshape((NA, 2 * NA, K3, K3, K3), order="F")


def loadHistograms(K2, K3, file, load_c1=True):
    """datagen.jl:432-475: (c2_full, c2, c3_full, c3[, c1_full]).  c3_full is the
    (doubled) histogram clamped to >= 1e-30, c3 its first K3 shells divided by
    their total; c2 likewise with K2."""
    _, c2_full, c3_full, c1_full = read_histograms(file)
    NA, NB = c3_full.shape[0], c3_full.shape[1]
    if NA == NB:   # doubling old histograms (:445-458)
        Kf = c3_full.shape[2]
        new = np.zeros((NA, 2 * NB, Kf, Kf, Kf))
        for k1 in range(Kf):
            for k2 in range(k1 + 1):
                for k3 in range(k2 + 1):
                    new[:, :NB, k3, k2, k1] = c3_full[:, :, k3, k2, k1]
                    new[:, NB:, k3, k2, k1] = c3_full[:, :, k3, k2, k1]
        c3_full = new
    c1_full = np.maximum(c1_full, 1e-30)
    c2_full = np.maximum(c2_full, 1e-30)
    c3 = load_c3(c3_full, K3)
    c3_full = np.maximum(c3_full, 1e-30)
    c2 = c2_full[:, :K2, :K2]
    c2 = c2 / np.abs(c2).sum()
    if load_c1:
        return c2_full, c2, c3_full, c3, c1_full
    return c2_full, c2, c3_full, c3


# ---- start structure -----------------------------------------------------------

def randomStartStructure(volume, K, L, seed):
    """determination.jl:4-17: the l-blocks (even l >= 2) of shells 1..K rotated by
    independent Haar-random SO(2l+1) matrices (random_rotation, utilities.jl:165-176,
    drawn from numpy's Philox keyed by `seed`)."""
    from .workload import haar_rotation, rng, umat
    from . import seanindex
    g = rng(seed)
    c = volume.coeff.copy()
    for l in range(2, L + 1, 2):
        U = umat(l)
        rot = U.conj().T @ haar_rotation(2 * l + 1, g).astype(np.complex128) @ U
        idx = [seanindex(l, m, volume.LMAX) - 1 for m in range(-l, l + 1)]
        c[:K, idx] = c[:K, idx] @ rot.T
    return SphericalHarmonicsVolume(c, volume.LMAX, volume.KMAX, volume.rmax)


# ---- one annealing run ---------------------------------------------------------

class AnnealRun:
    """rotate_all_at_once's state for one resolution L, with the structure on
    the device.  Attributes mirror the reference's state dict: i (steps taken),
    E, T, stepsizes {l: value}, naccepted; stopped once the stopping rule holds."""

    def __init__(self, state, basis, c3ref, K3_range=None, initial_stepsize=math.pi / 4, stepsizefactor=1.01,
                 initial_temperature_factor=0.1, temperature_decay=0.99998, seed=0, ntemp=50):
        self.dstate, self.basis = state, basis
        self.K3_range = K3_range if K3_range is not None else range(1, basis.K + 1)
        self.ref = basis.ref(c3ref, self.K3_range)
        self.params = AnnealParams(0, ntemp, initial_stepsize, stepsizefactor, initial_temperature_factor,
                                   temperature_decay, seed)
        self.status = AnnealStatus()
        self.started = False

    # state dict views
    @property
    def i(self):
        return int(self.status.step)

    @property
    def E(self):
        return self.status.E

    @property
    def T(self):
        return self.status.T

    @property
    def naccepted(self):
        return int(self.status.naccepted)

    @property
    def stopped(self):
        return bool(self.status.stopped)

    @property
    def stepsizes(self):
        return {2 * (q + 1): self.status.stepsizes[q] for q in range(self.status.nl)}

    def begin(self):
        """Initial energy, step sizes and temperature (determination.jl:291-312)."""
        check(lib().tp_anneal_begin(self.basis.ctx.h, self.basis.h, self.ref, self.dstate.h, C.byref(self.params),
                                    C.byref(self.status)), "tp_anneal_begin")
        self.started = True
        return self

    def steps(self, n, min_stepsize=0.0):
        """Up to n further steps (fewer once the stopping rule holds): traces of the
        steps taken, E after each step, proposal energies, acceptances."""
        if not self.started:
            raise RuntimeError("begin() or load() first")
        E = np.zeros(max(n, 1))
        En = np.zeros(max(n, 1))
        acc = np.zeros(max(n, 1), np.int32)
        done = C.c_int32()
        check(lib().tp_anneal_steps(self.basis.ctx.h, self.basis.h, self.ref, self.dstate.h, C.byref(self.params),
                                    C.byref(self.status), int(n), float(min_stepsize), E.ctypes.data,
                                    En.ctypes.data, acc.ctypes.data, C.byref(done)), "tp_anneal_steps")
        k = done.value
        return dict(E=E[:k], E_new=En[:k], accepted=acc[:k])

    def coefficients(self):
        return self.dstate.read()

    def save(self, path, min_stepsize=0.0):
        """saveState (determination.jl:264-278): parameters, status and structure."""
        c = np.ascontiguousarray(self.coefficients(), dtype=np.complex128)
        b = self.basis
        info = CheckpointInfo(self.params, self.status, float(min_stepsize), b.lam, b.dq, b.L, b.LMAX, b.N, b.K,
                              self.K3_range.start, self.K3_range.stop - 1, c.shape[0], c.shape[1])
        check(lib().tp_checkpoint_save(os.fsencode(path), C.byref(info), c.ctypes.data), "tp_checkpoint_save")

    @classmethod
    def load(cls, path, basis, c3ref, template=None):
        """Resume from a checkpoint: the structure goes back to the device and the
        next step continues the saved run's Philox streams exactly."""
        info = read_checkpoint_info(path)
        b = basis
        if (info.L, info.LMAX, info.N, info.K) != (b.L, b.LMAX, b.N, b.K) or info.lam != b.lam:
            raise ValueError(f"checkpoint was made on L={info.L} LMAX={info.LMAX} N={info.N} K={info.K} "
                             f"lambda={info.lam}, not on this basis")
        coeff = np.zeros((info.nshell, info.ld), np.complex128)
        check(lib().tp_checkpoint_load(os.fsencode(path), C.byref(info), coeff.ctypes.data, coeff.size * 2),
              "tp_checkpoint_load")
        vol = SphericalHarmonicsVolume(coeff, b.LMAX, info.nshell, template.rmax if template else 1.0) \
            if info.ld == b.LMAX * b.LMAX else coeff
        st = DeviceState(vol, b, info.nshell)
        p = info.params
        run = cls(st, b, c3ref, range(info.k3lo, info.k3hi + 1), p.initial_stepsize, p.stepsizefactor,
                  p.initial_temperature_factor, p.temperature_decay, p.seed, p.ntemp)
        run.status = info.status
        run.started = True
        run.min_stepsize = info.min_stepsize
        return run


def read_checkpoint_info(path):
    info = CheckpointInfo()
    check(lib().tp_checkpoint_load(os.fsencode(path), C.byref(info), None, 0), "tp_checkpoint_load")
    return info


def rotate_all_at_once(state, basis, c3ref, nsteps=None, initial_stepsize=math.pi / 4, stepsizefactor=1.01,
                       initial_temperature_factor=0.1, temperature_decay=0.99998, seed=0, ntemp=50, K3_range=None,
                       min_stepsize=minstepsize, checkpoint=None, checkpoint_every=CHECKPOINT_EVERY, log=None):
    """determination.jl:281-338 on the device: steps until the step size of the
    largest l is at or below min_stepsize (nsteps=None; the reference's loop) or
    nsteps steps, a checkpoint to `checkpoint` every checkpoint_every steps, and
    one log line per step to `log` (a text file object) as evaluate_step writes
    (:510).  Returns the traces and the final run (AnnealRun)."""
    run = AnnealRun(state, basis, c3ref, K3_range, initial_stepsize, stepsizefactor, initial_temperature_factor,
                    temperature_decay, seed, ntemp).begin()
    return continue_run(run, nsteps, min_stepsize if nsteps is None else 0.0, checkpoint, checkpoint_every, log)


def continue_run(run, nsteps=None, min_stepsize=minstepsize, checkpoint=None, checkpoint_every=CHECKPOINT_EVERY,
                 log=None, reference_energy=0.0):
    """Steps of `run` in pieces of checkpoint_every, until nsteps more steps are
    done or the stopping rule holds."""
    traces = {"E": [], "E_new": [], "accepted": []}
    left = nsteps
    while not run.stopped and (left is None or left > 0):
        n = checkpoint_every - (run.i % checkpoint_every)
        if left is not None:
            n = min(n, left)
        T, Eprev, i0 = run.T, run.E, run.i
        tr = run.steps(n, min_stepsize)
        k = len(tr["E"])
        for key in traces:
            traces[key].append(tr[key])
        if log is not None:
            # evaluate_step's line (determination.jl:510): i, sign, E_new, E, delta_E,
            # reference energy, E - reference, T (T *= decay before each step, :328)
            for j in range(k):
                T *= run.params.temperature_decay
                dE = tr["E_new"][j] - Eprev
                sign = (1.0 if dE < 0 else -1.0) if tr["accepted"][j] else 0.0
                Eprev = tr["E"][j]
                log.write(f"{i0 + j + 1}\t{sign}\t{tr['E_new'][j]}\t{Eprev}\t{dE}\t{reference_energy}\t"
                          f"{Eprev - reference_energy}\t{T}\n")
        if left is not None:
            left -= k
        # every checkpoint_every steps (regular_action), at the stop, and where this call ends
        if checkpoint is not None and (run.i % checkpoint_every == 0 or run.stopped or left == 0):
            run.save(checkpoint, min_stepsize)
        if k == 0:
            break
    out = {key: (np.concatenate(v) if v else np.zeros(0)) for key, v in traces.items()}
    out.update(E0=run.status.E0, T0=run.status.T0, E_final=run.E, T_final=run.T, steps=run.i,
               stopped=run.stopped, run=run)
    return out


# ---- rotation_search -------------------------------------------------------------

def rotation_search(params, workdir=".", nsteps=None, K2_dim=None):
    """determination.jl:154-229.  params (the reference's keys): "histograms"
    (a write_histograms file), "L", "LMAX", "N", "K2_range", "K3_range",
    "qmax", "lambda", "initial_stepsize", "stepsizefactor",
    "initial_temperature_factor", "temperature_decay", "measure" ("Bayes"),
    optional "reference_pdb_path" and "seed" (the Philox key of the start
    structure and of the annealing).  Files in workdir: params.json,
    state.ckpt (the checkpoint, every 500 steps and at the end), state.json
    (the run's progress: "running" / "finished_structure") and det.out (the
    per-step log).  A second call continues a stopped process from
    state.ckpt, and returns at once for a finished one.  nsteps bounds the
    steps of this call (None: until the stopping rule)."""
    os.makedirs(workdir, exist_ok=True)
    sfile, cfile = os.path.join(workdir, "state.json"), os.path.join(workdir, "state.ckpt")
    state = json.load(open(sfile)) if os.path.exists(sfile) else {"newRun": True}
    if state.get("state") == "finished_structure":
        return params, state
    if params.get("measure", "Bayes") != "Bayes":
        raise ValueError("only measure = \"Bayes\" has an energy (correlations.jl:210)")
    K3max = max(params["K3_range"])
    K2max = max(params["K2_range"])
    c2ref_full, c2ref, c3ref_full, c3ref = loadHistograms(K3max, K3max, params["histograms"], False)
    L, LMAX, N, lam = params["L"], params["LMAX"], params["N"], params.get("lambda", 0.0)
    seed = int(params.get("seed", 0))
    log = open(os.path.join(workdir, "det.out"), "w+" if state["newRun"] else "a")
    try:
        if state["newRun"]:
            state["newRun"] = False
            state["K"] = K3max
            json.dump(params, open(os.path.join(workdir, "params.json"), "w"))
            if params.get("reference_pdb_path"):
                from .structure import createSphericalHarmonicsStructure
                _, _, ref_int = createSphericalHarmonicsStructure(params["reference_pdb_path"], LMAX, K2max,
                                                                  math.pi / (params["qmax"] / K2max))
                np.save(os.path.join(workdir, "reference_intensity.npy"), ref_int.coeff)
            start = retrieveSolution(c2ref_full / np.abs(c2ref_full).sum(), L, LMAX, params["K2_range"],
                                     params["qmax"], lam)
            start = randomStartStructure(start, start.KMAX, L, seed)
            basis = HIP_calculate_basis(L, LMAX, N, K3max, lam, start.dq())
            run = AnnealRun(DeviceState(start, basis), basis, c3ref, range(1, K3max + 1),
                            params["initial_stepsize"], params["stepsizefactor"],
                            params["initial_temperature_factor"], params["temperature_decay"], seed).begin()
            state["reference_energy"] = _reference_energy(workdir, basis, c3ref, K3max)
            state["state"] = "running"
            state["rmax"] = start.rmax
            log.write(f"#Starting structure determination with the following parameters:\n#{params}\n")
            run.save(cfile, minstepsize)
        else:
            info = read_checkpoint_info(cfile)
            basis = HIP_calculate_basis(L, LMAX, N, K3max, lam, info.dq)
            run = AnnealRun.load(cfile, basis, c3ref)
        json.dump(state, open(sfile, "w"))
        res = continue_run(run, nsteps, minstepsize, cfile, CHECKPOINT_EVERY, log, state["reference_energy"])
        state.update(i=run.i, E=run.E, T=run.T, stepsizes={str(k): v for k, v in run.stepsizes.items()})
        if run.stopped:
            state["state"] = "finished_structure"
        json.dump(state, open(sfile, "w"))
        state["result"] = SphericalHarmonicsVolume(run.coefficients(), LMAX, run.dstate.nshell, state["rmax"]) \
            if run.dstate.template.shape[1] == LMAX * LMAX else run.coefficients()
        state["traces"] = res
        return params, state
    finally:
        log.close()


def _reference_energy(workdir, basis, c3ref, K3max):
    from . import energy
    p = os.path.join(workdir, "reference_intensity.npy")
    if not os.path.exists(p):
        return 0.0
    c = np.load(p, allow_pickle=False)
    return energy(c, basis, c3ref, range(1, K3max + 1))
