"""Gradient-guided annealing moves.

The reference's move (get_MC_step, determination.jl:232-241) rotates every
even-l block of all shells by an element of SO(2l+1) in the real basis, which
leaves the two-photon invariants unchanged; it picks the rotation at random.
Here the rotation follows the energy's steepest descent on that same group:
with x_k = Umat(l) c_k,l (real) and dE/dx_k = Re(Umat(l) gamma_k,l) from the
device gradient (tp_energy_grad), the Riemannian gradient on SO(2l+1) is the
skew part Omega_l of sum_k (dE/dx_k) x_k^T, and a move is R_l = exp(-eta
Omega_l) for every l at once.  Each step scores a ladder of step lengths eta
in one batched evaluation (the basis's precision: FP16 hi/lo in the intended
use), takes the best, and accepts it by Metropolis on its exact FP64 energy
(evaluate_step, determination.jl:491-515, with T *= decay per step as in
rotate_all_at_once :328).  The rotations are small dense matrices built on
the host (as the reference builds its Haar rotations on the host); every c3,
energy and gradient runs on the device.
"""
import time

import numpy as np

from . import PREC_FP64, energy, energy_batch, energy_grad, seanindex
from .workload import umat


class GradientMoves:
    """Per-l real-basis transforms and the rotation algebra of the moves."""

    def __init__(self, basis):
        self.basis = basis
        self.ls = list(range(2, basis.L + 1, 2))
        self.U = {l: umat(l) for l in self.ls}
        self.idx = {l: np.array([seanindex(l, m, basis.LMAX) - 1 for m in range(-l, l + 1)]) for l in self.ls}

    def riemannian(self, c, gamma):
        """{l: Omega_l}, the skew-symmetric steepest-ascent direction of E on SO(2l+1)."""
        K = gamma.shape[0]
        out = {}
        for l in self.ls:
            U, ix = self.U[l], self.idx[l]
            x = (c[:K, ix] @ U.T).real
            gx = (gamma[:, ix] @ U.T).real
            M = gx.T @ x
            out[l] = 0.5 * (M - M.T)
        return out

    @staticmethod
    def norm(om):
        return float(np.sqrt(sum((A * A).sum() for A in om.values())))

    def _block_rotations(self, om, etas):
        """{l: (S, n, n)} complex-basis rotations exp(-eta Omega_l) for all etas:
        with i Omega_l = V diag(w) V^H (Hermitian) and W = Umat(l)^H V, the
        rotation is W diag(exp(i eta w)) W^H (one batched product per l)."""
        etas = np.asarray(etas, dtype=np.float64)
        out = {}
        for l, A in om.items():
            w, V = np.linalg.eigh(1j * A)
            W = self.U[l].conj().T @ V
            ph = np.exp(1j * etas[:, None] * w[None, :])
            out[l] = (W[None, :, :] * ph[:, None, :]) @ W.conj().T
        return out

    def candidates(self, c, om, etas, nshell=None):
        """(S, nshell, LMAX^2): the coefficients moved by exp(-eta Omega_l) on
        every l, one slice per eta.  Only the entries the energy reads (even
        l <= L) are set; the others are left unset (the library never reads
        them).  Use moved() for a complete state."""
        nshell = c.shape[0] if nshell is None else nshell
        rots = self._block_rotations(om, etas)
        out = np.empty((len(etas), nshell, c.shape[1]), np.complex128)
        out[:, :, 0] = c[None, :nshell, 0]   # l = 0 (seanindex 1) is never rotated
        for l, rot in rots.items():
            ix = self.idx[l]
            out[:, :, ix] = c[None, :nshell, ix] @ np.swapaxes(rot, 1, 2)
        return out

    def moved(self, c, om, eta):
        """The complete coefficient array (every shell and entry) moved by exp(-eta Omega_l)."""
        rots = self._block_rotations(om, [eta])
        out = c.copy()
        for l, rot in rots.items():
            ix = self.idx[l]
            out[:, ix] = c[:, ix] @ rot[0].T
        return out

    def rotations(self, om, etas):
        """[{l: exp(-eta Omega_l)} for eta in etas] (real basis)."""
        out = [dict() for _ in etas]
        for l, A in om.items():
            w, V = np.linalg.eigh(1j * A)
            Vh = V.conj().T
            for s, eta in enumerate(etas):
                out[s][l] = ((V * np.exp(1j * eta * w)) @ Vh).real
        return out

    def apply(self, c, rots):
        """Coefficients with every shell's l-block rotated: x <- R x in the real basis."""
        out = c.copy()
        for l, R in rots.items():
            U, ix = self.U[l], self.idx[l]
            rot = U.conj().T @ R.astype(np.complex128) @ U
            out[:, ix] = c[:, ix] @ rot.T
        return out


def _energy64(c, basis, c3ref, K3_range):
    """FP64 energy whatever the basis's precision (the acceptance energy)."""
    prec = basis.precision()
    if prec == PREC_FP64:
        return energy(c, basis, c3ref, K3_range)
    basis.set_precision(PREC_FP64)
    try:
        return energy(c, basis, c3ref, K3_range)
    finally:
        basis.set_precision(prec)


def gradient_anneal(coeff, basis, c3ref, nsteps, ncand=16, T0=0.0, decay=1.0, seed=0, K3_range=None,
                    target=None, time_limit=None):
    """Anneal `coeff` (KMAX x LMAX^2 complex, all shells rotated) with gradient
    moves.  Per step: the FP64 energy gradient of the current state (only after
    it changed), ncand step lengths along exp(-eta Omega_l) scored in one
    batched evaluation in the basis's current precision, and the best one
    accepted by Metropolis on its FP64 energy.  The ladder of step lengths is
    re-centred on the accepted one (a quarter of it after a rejection).  Stops
    after nsteps, when E <= target, or after time_limit seconds.  Returns the
    final coefficients and energy, the (seconds, E) history (history[0] =
    start) and the step lengths eta |Omega|."""
    gm = GradientMoves(basis)
    rng = np.random.default_rng(seed)
    c = np.array(coeff.coeff if hasattr(coeff, "coeff") else coeff, dtype=np.complex128, copy=True)
    t0 = time.perf_counter()
    E, g = energy_grad(c, basis, c3ref, K3_range)
    hist, steps, accepted = [(0.0, E)], [], 0
    scale = None
    ladder = 2.0 ** ((np.arange(ncand) - ncand // 2) / 2.0)
    bases, refs, rngs = [basis] * ncand, [c3ref] * ncand, [K3_range] * ncand
    T = T0
    for step in range(nsteps):
        if g is None:
            _, g = energy_grad(c, basis, c3ref, K3_range)
        om = gm.riemannian(c, g)
        gn = gm.norm(om)
        if gn == 0.0:
            break
        if scale is None:
            scale = 1.0 / gn
        etas = scale * ladder
        cands = gm.candidates(c, om, etas, basis.K)
        Ec = energy_batch(list(cands), bases, refs, rngs)
        k = int(np.argmin(Ec))
        En = _energy64(cands[k], basis, c3ref, K3_range)
        T *= decay
        dE = En - E
        acc = dE < 0.0 or (T > 0.0 and rng.uniform() < np.exp(-dE / T))
        if acc:
            c, E, g = gm.moved(c, om, etas[k]), En, None
            scale = etas[k]
            accepted += 1
        else:
            scale = etas[k] / 4.0
        steps.append(float(etas[k] * gn))
        now = time.perf_counter() - t0
        hist.append((now, E))
        if (target is not None and E <= target) or (time_limit is not None and now >= time_limit):
            break
    return {"coeff": c, "E": E, "history": hist, "step_lengths": steps, "accepted": accepted}


# ---- learned proposals along the gradient geodesic ----------------------------
#
# A proposal turns every l-block of all shells by exp(-theta_l Omega_l / |Omega_l|)
# (the Riemannian steepest descent of that block on SO(2l+1), unit length) with
# one angle theta_l per l.  A run keeps a step scale s (angle per unit of
# gradient norm); a proposal's angles are theta_l = s |Omega_l| rho_l c with a
# log-normal common factor c and a per-l profile rho (geometric mean 1).
# Without the network rho is log-normal about 1; with it rho is the network's
# prediction (MlpProposer, BF16 kernels) from one shell's state and the
# temperature, rho_l = tan(phi_l / 2) of the rotation angle phi_l of its l-th
# output quaternion, with log-normal noise.  Training pairs are (features of
# each shell, the profile of each of the best candidates of a step that pass
# the Metropolis test) logged from the method's own runs.  The reference's
# move, a rotation of each l-block in a Haar-random 2-plane
# (utilities.jl:180-190), fills the fallback slots.  All candidates of a step
# are scored in one batched evaluation; the best is accepted by Metropolis on
# its FP64 energy.

def shell_features(c, k, T, gnorm, moves):
    """The network input for shell k: its 190 real-basis coefficients (even
    l <= 18, Umat(l) c_k,l; scaled to norm 8), log10 T and log10 |Omega|."""
    parts = [np.array([c[k, 0].real])]
    for l in moves.ls:
        parts.append((moves.U[l] @ c[k, moves.idx[l]]).real)
    x = np.concatenate(parts)
    n = np.linalg.norm(x)
    out = np.zeros(192, np.float32)
    out[:x.size] = 8.0 * x / n if n > 0 else x
    out[190] = np.log10(max(T, 1e-12))
    out[191] = np.log10(max(gnorm, 1e-300))
    return out


def angles_of(Y, nl):
    """theta_l = 2 atan2(|(x, y, z)|, |w|) of the network's quaternion per l."""
    q = Y[:, :4 * nl].reshape(-1, nl, 4).astype(np.float64)
    return 2.0 * np.arctan2(np.linalg.norm(q[:, :, 1:], axis=2), np.abs(q[:, :, 0]))


def ratios_of(Y, nl):
    """rho_l = tan(phi_l / 2) = |(x, y, z)| / |w| of the network's quaternion per l."""
    q = Y[:, :4 * nl].reshape(-1, nl, 4).astype(np.float64)
    return np.linalg.norm(q[:, :, 1:], axis=2) / np.maximum(np.abs(q[:, :, 0]), 1e-300)


def ratio_labels(rho):
    """Training targets: the quaternion (1, rho, 0, 0) / sqrt(1 + rho^2) per l
    (rotation angle 2 atan rho)."""
    t = np.zeros((rho.shape[0], 36), np.float32)
    nl = rho.shape[1]
    h = 1.0 / np.sqrt(1.0 + rho * rho)
    t[:, 0:4 * nl:4] = h
    t[:, 1:4 * nl:4] = rho * h
    return t


def _gmean_normalised(r):
    """Rows of r divided by their geometric mean."""
    return r / np.exp(np.log(r).mean(axis=1, keepdims=True))


def angle_labels(theta):
    """Training targets: the quaternion (cos theta/2, sin theta/2, 0, 0) per l."""
    t = np.zeros((theta.shape[0], 36), np.float32)
    nl = theta.shape[1]
    t[:, 0:4 * nl:4] = np.cos(theta / 2)
    t[:, 1:4 * nl:4] = np.sin(theta / 2)
    return t


class GeodesicMoves(GradientMoves):
    """Rotations exp(-theta_l Omega_l / |Omega_l|) with per-candidate, per-l
    angles, and draws of the reference's random move."""

    def plane_rotations(self, stepsizes, rng, S):
        """S draws of the reference's move (get_MC_step, determination.jl:232-241;
        random_rotation_step, utilities.jl:180-190): per l a rotation by
        beta ~ N(0, stepsize_l) in the plane of a random orthonormal 2-frame of
        R^(2l+1), as complex-basis matrices {l: (S, n, n)}."""
        out = {}
        for li, l in enumerate(self.ls):
            n = 2 * l + 1
            g = rng.normal(size=(S, 2, n))
            u = g[:, 0] / np.linalg.norm(g[:, 0], axis=1, keepdims=True)
            v = g[:, 1] - (g[:, 1] * u).sum(1, keepdims=True) * u
            v /= np.linalg.norm(v, axis=1, keepdims=True)
            b = rng.normal(size=S) * stepsizes[li]
            R = np.eye(n)[None] + (np.cos(b) - 1.0)[:, None, None] * (u[:, :, None] * u[:, None, :] +
                                                                      v[:, :, None] * v[:, None, :]) \
                + np.sin(b)[:, None, None] * (v[:, :, None] * u[:, None, :] - u[:, :, None] * v[:, None, :])
            U = self.U[l]
            out[l] = U.conj().T[None] @ R.astype(np.complex128) @ U[None]
        return out

    def angle_rotations(self, om, theta):
        """Complex-basis rotations exp(-theta_l Omega_l / |Omega_l|), {l: (S, n, n)}
        (by the eigen-decomposition of i Omega_l / |Omega_l|)."""
        out = {}
        for li, l in enumerate(self.ls):
            A = om[l]
            nA = np.sqrt((A * A).sum())
            if nA == 0.0:
                out[l] = np.broadcast_to(np.eye(2 * l + 1, dtype=np.complex128), (theta.shape[0], 2 * l + 1, 2 * l + 1))
                continue
            w, V = np.linalg.eigh(1j * A / nA)
            W = self.U[l].conj().T @ V
            ph = np.exp(1j * theta[:, li][:, None] * w[None, :])
            out[l] = (W[None, :, :] * ph[:, None, :]) @ W.conj().T
        return out

    def generators(self, om):
        """(gen, norms): Omega_l / |Omega_l|_F (zero where Omega_l = 0), real
        row-major, concatenated over l as tp_batch_eval_rotations reads them,
        and the norms |Omega_l|_F."""
        parts, norms = [], []
        for l in self.ls:
            A = om[l]
            nA = float(np.sqrt((A * A).sum()))
            norms.append(nA)
            parts.append((A / nA if nA > 0.0 else np.zeros_like(A)).ravel())
        return np.concatenate(parts), np.array(norms)

    def geodesic_rotations(self, gen, theta):
        """Complex-basis rotations exp(-theta_l G_l), {l: (S, n, n)}, for the
        generators gen of generators() (expm_skew, the device's algorithm)."""
        out, off = {}, 0
        for li, l in enumerate(self.ls):
            n = 2 * l + 1
            G = gen[off:off + n * n].reshape(n, n)
            off += n * n
            R = expm_skew(-theta[:, li][:, None, None] * G[None])
            U = self.U[l]
            out[l] = U.conj().T[None] @ R.astype(np.complex128) @ U[None]
        return out

    def rotate_all(self, c, rots, nshell):
        """(S, nshell, LMAX^2): shells 1..nshell of c with every l-block turned by
        rots[l][s]; entries outside even l <= L copied."""
        S = next(iter(rots.values())).shape[0]
        out = np.broadcast_to(c[None, :nshell], (S, nshell, c.shape[1])).copy()
        for l, rot in rots.items():
            ix = self.idx[l]
            out[:, :, ix] = c[None, :nshell, ix] @ np.swapaxes(rot, 1, 2)
        return out


def expm_skew(A):
    """exp(A) of real skew-symmetric matrices (..., n, n): scaled by 2^-s to
    |A|_F <= 1/4 (s per matrix), the degree-12 Taylor polynomial by Horner's
    rule, squared s times (the algorithm of the device's k_rotation_batch)."""
    A = np.asarray(A, dtype=np.float64)
    shp = A.shape
    A = A.reshape((-1,) + shp[-2:])
    n = shp[-1]
    out = np.empty_like(A)
    I = np.eye(n)
    for i in range(A.shape[0]):
        a, s = float(np.sqrt((A[i] * A[i]).sum())), 0
        while a > 0.25 and s < 64:
            a *= 0.5
            s += 1
        B = A[i] * 2.0 ** -s
        X = I + B / 12.0
        for k in range(11, 0, -1):
            X = I + (B @ X) / k
        for _ in range(s):
            X = X @ X
        out[i] = X
    return out.reshape(shp)


class MoveLog:
    """Training pairs of the learned proposer, in preallocated arrays: X
    (n, 192) float32 features, T (n, 36) float32 targets (angle quaternions)."""

    def __init__(self, capacity):
        self.X = np.zeros((capacity, 192), np.float32)
        self.T = np.zeros((capacity, 36), np.float32)
        self.n = 0

    def add(self, feats, labels):
        """Every row of labels with every row of feats (all pairs), as far as room is left."""
        m = min(feats.shape[0] * labels.shape[0], self.X.shape[0] - self.n)
        if m <= 0:
            return
        X = np.broadcast_to(feats[None, :, :], (labels.shape[0],) + feats.shape).reshape(-1, 192)[:m]
        T = np.broadcast_to(labels[:, None, :], (labels.shape[0], feats.shape[0], 36)).reshape(-1, 36)[:m]
        self.X[self.n:self.n + m] = X
        self.T[self.n:self.n + m] = T
        self.n += m

    def full(self):
        return self.n >= self.X.shape[0]

    def arrays(self):
        return self.X[:self.n], self.T[:self.n]


def _plane_draws(nl, ls, stepsizes, rng, S, L):
    """S draws of the reference's move per l (get_MC_step, determination.jl:232-241;
    random_rotation_step, utilities.jl:180-190): (u, v, b) rows of
    tp_batch_eval_rotations, u, v a random orthonormal 2-frame of R^(2l+1)
    and b ~ N(0, stepsize_l)."""
    width = 2 * (2 * L + 1) + 1
    out = np.zeros((S, nl, width))
    for li, l in enumerate(ls):
        n = 2 * l + 1
        g = rng.normal(size=(S, 2, n))
        u = g[:, 0] / np.linalg.norm(g[:, 0], axis=1, keepdims=True)
        v = g[:, 1] - (g[:, 1] * u).sum(1, keepdims=True) * u
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        out[:, li, :n] = u
        out[:, li, n:2 * n] = v
        out[:, li, 2 * n] = rng.normal(size=S) * stepsizes[li]
    return out


def plane_rotation(u, v, b):
    """I + (cos b - 1)(u u^T + v v^T) + sin b (v u^T - u v^T) (utilities.jl:180-190)."""
    return np.eye(u.size) + (np.cos(b) - 1.0) * (np.outer(u, u) + np.outer(v, v)) + \
        np.sin(b) * (np.outer(v, u) - np.outer(u, v))


class _Run:
    """Host state of one run of learned_gradient_runs."""

    def __init__(self, c, ref, K, T, stepsize0, ls):
        self.c, self.ref, self.K, self.T = c, ref, K, T
        self.E, self.g, self.scale = None, None, None
        self.stepsizes = np.array([stepsize0 * (l - 1) for l in ls])
        self.accepted, self.wins, self.hist = 0, {"geodesic": 0, "plane": 0}, []
        self.done = False


def learned_gradient_runs(coeffs, basis, c3refs, nsteps, mlp=None, nlearn=48, nrand=16, noise=0.25, spread=0.7,
                          T0=0.0, decay=1.0, seed=0, targets=None, time_limit=None, log=None, log_top=8,
                          stepsizefactor=1.01, initial_stepsize=np.pi / 4, batch=None, record=None):
    """Concurrent annealing runs along the energy gradient, all candidates of
    a step scored in one batched evaluation.  Run g has the coefficients
    coeffs[g] (every shell is rotated), the histogram c3refs[g] of K_g shells
    (walked over K3_range 1:K_g) and per step nlearn geodesic candidates
    exp(-theta_l Omega_l / |Omega_l|) along its Riemannian steepest-descent
    direction Omega (tp_energy_grad in FP64, after the state changed) and
    nrand draws of the reference's random move (get_MC_step,
    determination.jl:232-241).  Geodesic angles: theta_l = s |Omega_l| rho_l
    exp(spread N) with the run's step scale s and one N(0, 1) draw per
    candidate; the profile rho (geometric mean 1) is the network's (mlp; slot
    j reads the features of shell j K_g / nlearn) times exp(noise N_l), or
    without it exp(spread N_l / 2).  s follows the geometric mean of
    theta / (s |Omega|) of the accepted geodesic moves and halves after a
    rejection; the plane moves' step sizes grow and shrink by stepsizefactor
    as in the reference (determination.jl:476-489).  All runs' candidates are
    turned and scored on the device (tp_batch_eval_rotations, the basis's
    mixed precision); the best of each run is accepted by Metropolis on its
    FP64 energy with T *= decay per step (evaluate_step,
    determination.jl:491-515) and handed back to the batch
    (tp_batch_set_state).  log (a MoveLog) receives, for each of the log_top
    best geodesic candidates that pass the Metropolis test on their own, the
    features of every shell of its run with that candidate's profile.  A run
    stops at targets[g] (energy); all stop at time_limit seconds.  batch: an
    AnnealBatch of these runs with nlearn + nrand slots (made here, in FP16
    hi/lo, when not given).  record (a list) receives per step and run the
    generator norms, step scale, geodesic angles, candidate energies and the
    best slot.  Returns per run {"coeff", "E", "history" [(t, E)],
    "accepted", "wins"} and the step count."""
    from . import AnnealBatch, HistogramRef, PREC_FP16X3
    gm = GeodesicMoves(basis)
    c3refs = [x if isinstance(x, HistogramRef) else basis.histogram(x, range(1, np.asarray(x).shape[2] + 1))
              for x in c3refs]
    rng = np.random.default_rng(seed)
    G, nl, L = len(coeffs), len(gm.ls), basis.L
    ncand = nlearn + nrand
    if ncand < 3:
        raise ValueError("need nlearn + nrand >= 3")
    cs, nrows = [], []
    for x in coeffs:
        a = np.asarray(x.coeff if hasattr(x, "coeff") else x, dtype=np.complex128)
        nrows.append(a.shape[0])
        c = np.zeros((max(a.shape[0], basis.K), a.shape[1]), np.complex128)   # shells past K_g stay zero
        c[:a.shape[0]] = a
        cs.append(c)
    if batch is None:
        prec = basis.precision()
        basis.set_precision(PREC_FP16X3)
        try:
            batch = AnnealBatch(basis, list(c3refs), cs, [seed + g for g in range(G)], nprop=ncand)
        finally:
            basis.set_precision(prec)
    B = batch
    if B.nprop != ncand or B.ngroups != G:
        raise ValueError("batch must hold the runs with nlearn + nrand slots each")
    Ks = list(B.Ks)
    runs = [_Run(cs[g], c3refs[g], Ks[g], T0, initial_stepsize, gm.ls) for g in range(G)]
    gsz = sum((2 * l + 1) ** 2 for l in gm.ls)
    width = 2 * (2 * L + 1) + 1
    gen = np.zeros((G, gsz))
    theta = np.zeros((G * ncand, nl))
    uvb = np.zeros((G * ncand, nl, width))
    norms = np.zeros((G, nl))
    t0 = time.perf_counter()
    for r in runs:
        r.E, r.g = energy_grad(r.c, basis, r.ref, range(1, r.K + 1))
        r.hist.append((0.0, r.E))
    for g, r in enumerate(runs):
        B.set_state(g, r.c, r.E)
    step = 0
    for step in range(1, nsteps + 1):
        feats = [None] * G
        for g, r in enumerate(runs):
            if r.g is None:
                _, r.g = energy_grad(r.c, basis, r.ref, range(1, r.K + 1))
            gen[g], norms[g] = gm.generators(gm.riemannian(r.c, r.g))
            if r.scale is None:
                gn = float(np.sqrt((norms[g] ** 2).sum()))
                r.scale = 1.0 / gn if gn > 0 else 0.0
            if mlp is not None or log is not None:
                gn = float(np.sqrt((norms[g] ** 2).sum()))
                feats[g] = np.stack([shell_features(r.c, k, r.T, gn, gm) for k in range(r.K)])
        if nlearn:
            if mlp is not None:
                X = np.concatenate([feats[g][[j * runs[g].K // nlearn for j in range(nlearn)]] for g in range(G)])
                rho = _gmean_normalised(np.maximum(ratios_of(mlp.forward(X), nl), 1e-3)) * \
                    np.exp(noise * rng.normal(size=(G * nlearn, nl)))
            else:
                rho = np.exp(0.5 * spread * rng.normal(size=(G * nlearn, nl)))
            common = np.exp(spread * rng.normal(size=(G * nlearn, 1)))
            base = np.repeat(np.array([r.scale for r in runs])[:, None] * norms, nlearn, axis=0)
            theta.reshape(G, ncand, nl)[:, :nlearn] = (base * rho * common).reshape(G, nlearn, nl)
        if nrand:
            for g, r in enumerate(runs):
                uvb[g * ncand + nlearn:(g + 1) * ncand] = _plane_draws(nl, gm.ls, r.stepsizes, rng, nrand, L)
        Ec = B.eval_rotations(nlearn, gen, theta, uvb).reshape(G, ncand)
        now = time.perf_counter() - t0
        for g, r in enumerate(runs):
            if r.done:
                continue
            r.T *= decay
            nz = norms[g] > 0
            if log is not None and nlearn and not log.full() and nz.all() and r.scale > 0:
                d = Ec[g, :nlearn] - r.E
                ok = d < 0.0
                if r.T > 0.0:
                    ok |= rng.uniform(size=nlearn) < np.exp(-np.maximum(d, 0.0) / r.T)
                top = [j for j in np.argsort(d, kind="stable")[:log_top] if ok[j]]
                if top:
                    ratio = theta[g * ncand + np.array(top)] / (r.scale * norms[g])[None, :]
                    log.add(feats[g], ratio_labels(_gmean_normalised(ratio)))
            k = int(np.argmin(Ec[g]))
            e = g * ncand + k
            if record is not None:
                record.append({"step": step, "run": g, "norms": norms[g].copy(), "scale": r.scale, "E": r.E,
                               "theta": theta[g * ncand:g * ncand + nlearn].copy(), "Ec": Ec[g].copy(), "best": k})
            if k < nlearn:
                rots = gm.geodesic_rotations(gen[g], theta[e:e + 1])
            else:
                rots = {}
                for li, l in enumerate(gm.ls):
                    n = 2 * l + 1
                    q = uvb[e, li]
                    R = plane_rotation(q[:n], q[n:2 * n], q[2 * n])
                    rots[l] = (gm.U[l].conj().T @ R.astype(np.complex128) @ gm.U[l])[None]
            cand = gm.rotate_all(r.c, rots, r.c.shape[0])[0]
            En = _energy64(cand, basis, r.ref, range(1, r.K + 1))
            dE = En - r.E
            if dE < 0.0 or (r.T > 0.0 and rng.uniform() < np.exp(-dE / r.T)):
                r.accepted += 1
                r.c, r.E, r.g = cand, En, None
                B.set_state(g, r.c, r.E)
                if k < nlearn:
                    r.wins["geodesic"] += 1
                    if nz.any():
                        r.scale = float(np.exp(np.log(theta[e][nz] / norms[g][nz]).mean()))
                else:
                    r.wins["plane"] += 1
                    r.stepsizes *= stepsizefactor
            else:
                r.scale /= 2.0
                r.stepsizes /= stepsizefactor
            r.hist.append((time.perf_counter() - t0, r.E))
            if targets is not None and targets[g] is not None and r.E <= targets[g]:
                r.done = True
        if all(r.done for r in runs) or (time_limit is not None and now >= time_limit):
            break
    return [{"coeff": r.c[:nrows[g]], "E": r.E, "history": r.hist, "accepted": r.accepted, "wins": r.wins}
            for g, r in enumerate(runs)], step


def learned_gradient_anneal(coeff, basis, c3ref, nsteps, mlp=None, target=None, **kw):
    """One run of learned_gradient_runs (its keyword arguments), as
    gradient_anneal returns it plus the moves won by each kind."""
    res, _ = learned_gradient_runs([coeff], basis, [c3ref], nsteps, mlp=mlp,
                                   targets=None if target is None else [target], **kw)
    return res[0]
