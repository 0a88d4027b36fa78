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
