"""ctypes view of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the cpu_baseline leg of bench.py -- never by the product package.  The C file
tp_oracle.c documents which reference lines each function restates.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

d_p = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
i64_p = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
i32_p = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
u32_p = np.ctypeslib.ndpointer(dtype=np.uint32, flags="C_CONTIGUOUS")


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = C.CDLL(_LIB)
        L.or_seanindex.restype = C.c_int64
        L.or_seanindex.argtypes = [C.c_int, C.c_int, C.c_int]
        L.or_tripletFactor.restype = C.c_double
        L.or_tripletFactor.argtypes = [C.c_int] * 3
        L.or_wigner3j.restype = C.c_double
        L.or_wigner3j.argtypes = [C.c_int] * 6
        L.or_sphPlm.restype = C.c_double
        L.or_sphPlm.argtypes = [C.c_int, C.c_int, C.c_double]
        L.or_mlp_nparam.restype = C.c_int
        L.or_mlp_loss_grad.restype = C.c_double
        L.or_mlp_loss_grad.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]
        L.or_basis_create.restype = C.c_void_p
        L.or_basis_create.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int]
        L.or_basis_free.argtypes = [C.c_void_p]
        L.or_basis_info.argtypes = [C.c_void_p, i64_p]
        L.or_basis_tables.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.or_pacombos.argtypes = [C.c_void_p, i64_p]
        L.or_triple_products.argtypes = [C.c_void_p, d_p, C.c_int64, d_p, C.c_int]
        L.or_c3_dense.argtypes = [C.c_void_p, d_p, C.c_int64, C.c_int64, C.c_int64, d_p, C.c_int]
        L.or_factor_F.restype = C.c_int
        L.or_factor_F.argtypes = [C.c_void_p, d_p]
        L.or_c3_factored.restype = C.c_int
        L.or_c3_factored.argtypes = [C.c_void_p, d_p, C.c_int64, d_p, C.c_int]
        L.or_c3_slice.argtypes = [d_p, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                  C.c_double, C.c_double, d_p]
        L.or_energy_grad.restype = C.c_double
        L.or_energy_grad.argtypes = [C.c_void_p, d_p, C.c_int64, d_p, C.c_int, C.c_int, C.c_int, d_p, C.c_int]
        L.or_angle_between.restype = C.c_double
        L.or_angle_between.argtypes = [d_p, d_p]
        L.or_hist_pictures.restype = None
        L.or_hist_pictures.argtypes = [d_p, C.POINTER(C.c_int64), C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                       d_p, d_p, d_p, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                       C.POINTER(C.c_uint64)]
        L.or_energy.restype = C.c_double
        L.or_energy.argtypes = [d_p, C.c_int, C.c_int64, d_p, C.c_int, C.c_int, C.c_int]
        L.or_umat.argtypes = [C.c_int, d_p]
        L.or_haar_from_gaussian.restype = C.c_int
        L.or_haar_from_gaussian.argtypes = [C.c_int, d_p, d_p]
        L.or_rotation_step_matrix.argtypes = [C.c_int, d_p, C.c_double, d_p]
        L.or_frame_rotation_matrix.argtypes = [C.c_int, d_p, d_p, C.c_double, d_p]
        L.or_apply_real_rotation.argtypes = [C.c_int, d_p, d_p, C.c_int64, C.c_int, C.c_int]
        L.or_philox4x32_10.argtypes = [u32_p, u32_p, u32_p]
        L.or_normal.restype = C.c_double
        L.or_normal.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
        L.or_uniform.restype = C.c_double
        L.or_uniform.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32]
        L.or_mc_propose.argtypes = [C.c_uint64, C.c_uint32, C.c_uint32, C.c_int, d_p, d_p, d_p, C.c_int64,
                                    C.c_int, C.c_int]
        L.or_anneal.restype = C.c_int
        L.or_anneal.argtypes = [C.c_void_p, d_p, d_p, C.c_int64, C.c_int, C.c_int, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_uint64, C.c_int, d_p, d_p, i32_p, d_p, C.c_int]
        L.or_anneal_ex.restype = C.c_int
        L.or_anneal_ex.argtypes = [C.c_void_p, d_p, d_p, C.c_int64, C.c_int, C.c_int, d_p, d_p, C.c_int,
                                   C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, C.c_uint64, C.c_int,
                                   d_p, d_p, i32_p, C.c_int]
        L.or_load_c3.restype = None
        L.or_load_c3.argtypes = [d_p, C.c_int, C.c_int, C.c_int, C.c_int, d_p]
        L.or_d_work.restype = C.c_double
        L.or_d_work.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double]
        L.or_rotation_matrix.argtypes = [C.c_double, C.c_double, C.c_double, C.c_int, d_p]
        L.or_gauss_legendre.argtypes = [C.c_int, d_p, d_p]
        L.or_intensity_from_atoms.argtypes = [d_p, C.c_int, C.c_int, C.c_double, C.c_int, C.c_int, C.c_int,
                                              C.c_int, d_p, C.c_int]
        _lib = L
    return _lib


TAG_MOVE = 0x4D4F5645
TAG_ACC = 0x41434350
TAG_TEMP = 0x54454D50


def seanindex(l, m, LMAX):
    return lib().or_seanindex(l, m, LMAX)


def wigner3j(l1, m1, l2, m2, l3, m3):
    return lib().or_wigner3j(l1, m1, l2, m2, l3, m3)


def sphPlm(l, m, theta):
    return lib().or_sphPlm(l, m, theta)


def tripletFactor(k1, k2, k3):
    return lib().or_tripletFactor(k1, k2, k3)


class Basis:
    """calculate_basis (correlations.jl:87-153) restated in the oracle."""

    def __init__(self, L, LMAX, N, K, lam=0.0, dq=0.0, for_intensity=True, fp32_tables=False):
        self.L, self.LMAX, self.N, self.K, self.lam, self.dq = L, LMAX, N, K, lam, dq
        self.h = lib().or_basis_create(L, LMAX, N, K, lam, dq, int(for_intensity), int(fp32_tables))
        if not self.h:
            raise ValueError("bad basis parameters")
        info = np.zeros(5, np.int64)
        lib().or_basis_info(self.h, info)
        self.basislen, self.n_lt, self.klength, self.R, self.n_m = (int(v) for v in info)

    def __del__(self):
        if getattr(self, "h", None):
            lib().or_basis_free(self.h)
            self.h = None

    def tables(self):
        w = np.zeros(self.basislen)
        ind = np.zeros(9 * self.basislen, np.int64)
        lt = np.zeros(5 * self.n_lt, np.int64)
        lib().or_basis_tables(self.h, w.ctypes.data, ind.ctypes.data, lt.ctypes.data)
        return w, ind.reshape(self.basislen, 9), lt.reshape(self.n_lt, 5)

    def pacombos(self):
        out = np.zeros(9 * self.klength * self.n_lt, np.int64)
        lib().or_pacombos(self.h, out)
        return out.reshape(-1, 9)

    def triple_products(self, coeff, nthreads=0):
        coeff = _coeff_arr(coeff)
        As = np.zeros(self.klength * self.n_lt)
        lib().or_triple_products(self.h, coeff, coeff.shape[1] // 2, As, nthreads)
        return As.reshape(self.klength, self.n_lt)

    def c3_dense(self, coeff, col_lo=0, col_hi=None, nthreads=0):
        coeff = _coeff_arr(coeff)
        if col_hi is None:
            col_hi = self.klength
        out = np.zeros(self.R * (col_hi - col_lo))
        lib().or_c3_dense(self.h, coeff, coeff.shape[1] // 2, col_lo, col_hi, out, nthreads)
        return out.reshape(col_hi - col_lo, self.R).T   # R x ncols

    def F(self):
        F = np.zeros(self.R * self.n_lt)
        if lib().or_factor_F(self.h, F) != 0:
            raise ValueError("factored form needs lambda = 0")
        return F.reshape(self.n_lt, self.R).T

    def c3_factored(self, coeff, nthreads=0):
        coeff = _coeff_arr(coeff)
        out = np.zeros(self.R * self.klength)
        if lib().or_c3_factored(self.h, coeff, coeff.shape[1] // 2, out, nthreads) != 0:
            raise ValueError("factored form needs lambda = 0")
        return out.reshape(self.klength, self.R).T

    def energy_grad(self, coeff, c3ref, k3lo=1, k3hi=None, nthreads=0):
        """(E, gamma): the Bayes energy and gamma_k[i] = dE/dRe c + i dE/dIm c
        for the first K shells (complex, shaped like coeff[:K])."""
        c = _coeff_arr(np.asarray(coeff)[:self.K])
        cref = np.ascontiguousarray(np.asarray(c3ref, dtype=np.float64).ravel(order="F"))
        Kr = np.asarray(c3ref).shape[2]
        g = np.zeros_like(c)
        E = lib().or_energy_grad(self.h, c, c.shape[1] // 2, cref, Kr, k3lo, self.K if k3hi is None else k3hi, g,
                                 nthreads)
        return E, g.view(np.complex128)


def _coeff_arr(coeff):
    """complex (K, ld) -> float64 (K, 2*ld) contiguous; shells are rows."""
    a = np.ascontiguousarray(np.asarray(coeff, dtype=np.complex128))
    return a.view(np.float64).reshape(a.shape[0], 2 * a.shape[1]) if a.ndim == 2 else a


def c3_slice(coeff, N, L, LMAX, k1, k2, k3, dq=0.0, lam=0.0):
    c = _coeff_arr(coeff)
    out = np.zeros(2 * N * N)
    lib().or_c3_slice(c, c.shape[1] // 2, N, L, LMAX, k1, k2, k3, dq, lam, out)
    return out


def energy(c3raw, N, c3ref, Kr, k3lo, k3hi):
    """c3raw: R x ncols; c3ref: Julia array (N, 2N, Kr, Kr, Kr) stored column-major."""
    c3raw = np.asarray(c3raw, dtype=np.float64)
    craw = np.ascontiguousarray(c3raw.T).ravel()
    cref = np.ascontiguousarray(np.asarray(c3ref, dtype=np.float64).ravel(order="F"))
    return lib().or_energy(craw, N, c3raw.shape[1], cref, Kr, k3lo, k3hi)


def angle_between(p1, p2):
    return lib().or_angle_between(np.ascontiguousarray(p1, dtype=np.float64), np.ascontiguousarray(p2, dtype=np.float64))


def hist_pictures(photons, offsets, dq, N, K2, K3):
    """histogramCorrelationsInPicture_alltoall over pictures: photons (n, 3),
    picture i = photons[offsets[i]:offsets[i+1]].  Returns c1 (K2,), c2 (N, K2, K2),
    c3 (N, 2N, K3, K3, K3) indexed as the reference ([ai, k2, k1], [ai, bi, k3, k2, k1]
    zero-based) and the matching integer counts n1, n2, n3."""
    ph = np.ascontiguousarray(photons, dtype=np.float64)
    off = np.ascontiguousarray(offsets, dtype=np.int64)
    c1, c2, c3 = np.zeros(K2), np.zeros(N * K2 * K2), np.zeros(2 * N * N * K3 ** 3)
    n1 = np.zeros(K2, np.uint64)
    n2 = np.zeros(N * K2 * K2, np.uint64)
    n3 = np.zeros(2 * N * N * K3 ** 3, np.uint64)
    u64 = C.POINTER(C.c_uint64)
    lib().or_hist_pictures(ph, off.ctypes.data_as(C.POINTER(C.c_int64)), len(off) - 1, dq, N, K2, K3, c1, c2, c3,
                           n1.ctypes.data_as(u64), n2.ctypes.data_as(u64), n3.ctypes.data_as(u64))
    s2, s3 = (N, K2, K2), (N, 2 * N, K3, K3, K3)
    return (c1, c2.reshape(s2, order="F"), c3.reshape(s3, order="F"), n1, n2.reshape(s2, order="F"),
            n3.reshape(s3, order="F"))


def umat(l):
    n = 2 * l + 1
    out = np.zeros(2 * n * n)
    lib().or_umat(l, out)
    return out.view(np.complex128).reshape(n, n)


def haar_from_gaussian(A):
    A = np.ascontiguousarray(A, dtype=np.float64)
    n = A.shape[0]
    Cm = np.zeros((n, n))
    s = lib().or_haar_from_gaussian(n, A, Cm)
    return Cm, s


def rotation_step_matrix(Cm, beta):
    n = Cm.shape[0]
    out = np.zeros((n, n))
    lib().or_rotation_step_matrix(n, np.ascontiguousarray(Cm), beta, out)
    return out


def frame_rotation_matrix(u, v, beta):
    n = len(u)
    out = np.zeros((n, n))
    lib().or_frame_rotation_matrix(n, np.ascontiguousarray(u, dtype=np.float64),
                                   np.ascontiguousarray(v, dtype=np.float64), beta, out)
    return out


def apply_real_rotation(l, Rm, coeff, LMAX):
    c = np.array(coeff, dtype=np.complex128, copy=True)
    cf = c.view(np.float64).reshape(c.shape[0], -1)
    lib().or_apply_real_rotation(l, np.ascontiguousarray(Rm, dtype=np.float64), cf, c.shape[1], LMAX, c.shape[0])
    return c


def philox(ctr, key):
    out = np.zeros(4, np.uint32)
    lib().or_philox4x32_10(np.asarray(ctr, np.uint32), np.asarray(key, np.uint32), out)
    return out


def mc_propose(seed, step, tag, L, stepsizes, coeff, LMAX):
    c = np.ascontiguousarray(coeff, dtype=np.complex128)
    out = np.zeros_like(c)
    lib().or_mc_propose(seed, step, tag, L, np.ascontiguousarray(stepsizes, dtype=np.float64),
                        c.view(np.float64).reshape(c.shape[0], -1), out.view(np.float64).reshape(c.shape[0], -1),
                        c.shape[1], LMAX, c.shape[0])
    return out


def anneal(basis, c3ref, coeff, nsteps, initial_stepsize, stepsizefactor, initial_temperature_factor,
           temperature_decay, seed, ntemp=50, nthreads=0):
    c = np.array(coeff, dtype=np.complex128, copy=True)
    cref = np.ascontiguousarray(np.asarray(c3ref, dtype=np.float64).ravel(order="F"))
    E = np.zeros(nsteps)
    En = np.zeros(nsteps)
    acc = np.zeros(nsteps, np.int32)
    sc = np.zeros(4)
    rc = lib().or_anneal(basis.h, cref, c.view(np.float64).reshape(c.shape[0], -1), c.shape[1], c.shape[0], nsteps,
                         initial_stepsize, stepsizefactor, initial_temperature_factor, temperature_decay, seed,
                         ntemp, E, En, acc, sc, nthreads)
    if rc != 0:
        raise ValueError("anneal needs lambda = 0")
    return dict(E=E, E_new=En, accepted=acc, E0=sc[0], T0=sc[1], E_final=sc[2], T_final=sc[3], coeff=c)


class AnnealState:
    """rotate_all_at_once's state for the oracle's resumable driver (or_anneal_ex)."""

    def __init__(self, basis, c3ref, coeff, initial_stepsize, stepsizefactor, initial_temperature_factor,
                 temperature_decay, seed, ntemp=50, nthreads=0):
        self.basis = basis
        self.cref = np.ascontiguousarray(np.asarray(c3ref, dtype=np.float64).ravel(order="F"))
        self.coeff = np.array(coeff, dtype=np.complex128, copy=True)
        self.args = (initial_stepsize, stepsizefactor, initial_temperature_factor, temperature_decay, seed, ntemp)
        self.st = np.zeros(6)
        self.stepsizes = np.zeros(max(basis.L // 2, 1))
        self.nthreads = nthreads
        self._run(1, 0, 0.0)

    def _run(self, begin, n, min_stepsize):
        c = self.coeff
        E, En = np.zeros(max(n, 1)), np.zeros(max(n, 1))
        acc = np.zeros(max(n, 1), np.int32)
        k = lib().or_anneal_ex(self.basis.h, self.cref, c.view(np.float64).reshape(c.shape[0], -1), c.shape[1],
                               c.shape[0], begin, self.st, self.stepsizes, n, min_stepsize, *self.args, E, En, acc,
                               self.nthreads)
        if k < 0:
            raise ValueError("anneal needs lambda = 0")
        return dict(E=E[:k], E_new=En[:k], accepted=acc[:k])

    def steps(self, n, min_stepsize=0.0):
        return self._run(0, n, min_stepsize)

    @property
    def step(self):
        return int(self.st[0])

    @property
    def E(self):
        return self.st[1]

    @property
    def T(self):
        return self.st[2]


def load_c3(c3_full, K3):
    """loadHistograms' triplet histogram (datagen.jl:432-475), see or_load_c3."""
    a = np.asarray(c3_full, dtype=np.float64)
    NA, NB, Kf = a.shape[0], a.shape[1], a.shape[2]
    out = np.zeros(2 * NA * NA * K3 ** 3)
    lib().or_load_c3(np.ascontiguousarray(a.ravel(order="F")), NA, NB, Kf, K3, out)
    return out.reshape((NA, 2 * NA, K3, K3, K3), order="F")


def d_work(l, m1, m2, beta):
    return lib().or_d_work(l, m1, m2, beta)


def rotation_matrix(theta, phi, gamma, l):
    n = 2 * l + 1
    out = np.zeros(2 * n * n)
    lib().or_rotation_matrix(theta, phi, gamma, l, out)
    return out.view(np.complex128).reshape(n, n)


def gauss_legendre(n):
    x = np.zeros(n)
    w = np.zeros(n)
    lib().or_gauss_legendre(n, x, w)
    return x, w


def intensity_from_atoms(atoms, KMAX, dq, LMAX, lfill, ntheta, nphi, nthreads=0):
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    out = np.zeros((KMAX, LMAX * LMAX), np.complex128)
    lib().or_intensity_from_atoms(atoms, atoms.shape[0], KMAX, dq, LMAX, lfill, ntheta, nphi,
                                  out.view(np.float64).reshape(KMAX, -1), nthreads)
    return out


def mlp_nparam():
    return lib().or_mlp_nparam()


def mlp_loss_grad(params, X, T, emulate=True, want_grad=True):
    """Loss, gradient (FP64) and outputs of the proposer MLP (or_mlp_loss_grad)."""
    p = np.ascontiguousarray(params, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32)
    T = np.ascontiguousarray(T, dtype=np.float32)
    B = X.shape[0]
    g = np.zeros(p.size) if want_grad else None
    Y = np.zeros((B, 64))
    loss = lib().or_mlp_loss_grad(p.ctypes.data, B, X.ctypes.data, T.ctypes.data, int(emulate),
                                  g.ctypes.data if want_grad else None, Y.ctypes.data)
    return loss, g, Y
