"""Host-side mirror of the ThreePhotons.jl GPU backend (cuda.jl) over libtp.so.

Names, argument meaning and error behaviour follow the reference's Julia API
for this path:

  HIP_init()                         ~ CUDA_init            (cuda.jl:10-21)
  HIP_calculate_basis(...)           ~ CUDA_calculate_basis (cuda.jl:57-65)
  FullCorrelation_parallized(...)    (cuda.jl:74-108, correlations.jl:179-193)
  energy(...)                        (correlations.jl:196-226)
  rotate_all_at_once(...)            (determination.jl:281-338)

Arrays use the reference's index order with 0-based numpy indexing:
coefficients are a (KMAX, LMAX^2) complex array (row k-1 is intensity.coeff[k]
in seanindex order), c3 histograms are (N, 2N, K, K, K) arrays indexed
[a, b, k3, k2, k1].  Every call goes to the gfx950 library; there is no CPU
fallback here (the reference's own CPU BasisType is the fallback when HIP_init
fails).
"""
import ctypes as C

import numpy as np

from ._lib import PREC_BF16, PREC_FP16X3, PREC_FP64, AnnealParams, TPError, build, check, lib, track

__all__ = [
    "TPError", "build", "lib", "SphericalHarmonicsVolume", "HIP_init", "HipContext", "BasisTypeHip",
    "HIP_calculate_basis", "FullCorrelation_parallized", "energy", "DeviceState", "rotate_all_at_once",
    "intensity_from_atoms", "seanindex", "tripletFactor", "klength", "twoPhotons", "retrieveSolution",
    "HistogramRef",
]


def _digest(a):
    """64-bit hash of a contiguous array's bytes (xxh3; blake2b where xxhash is absent)."""
    try:
        import xxhash
        return xxhash.xxh3_64_intdigest(memoryview(a).cast("B"))
    except ImportError:
        import hashlib
        return hashlib.blake2b(memoryview(a).cast("B"), digest_size=8).hexdigest()


def seanindex(l, m, LMAX):
    """1-based position of (l, m) in a length-LMAX^2 vector (spherical_harmonics.jl:163-170)."""
    B = LMAX
    if m >= 0:
        return 1 + m * B - m * (m - 1) // 2 + (l - m)
    a = B - 1 + m
    return 1 + B * (B + 1) // 2 + a * (a + 1) // 2 + (l + m)


def tripletFactor(k1, k2, k3):
    """datagen.jl:166-174"""
    if k1 == k2 == k3:
        return 1.0
    if k1 == k2 or k2 == k3 or k1 == k3:
        return 3.0
    return 6.0


def klength(K):
    return K * (K + 1) * (K + 2) // 6


class SphericalHarmonicsVolume:
    """Per-shell complex SH coefficients (spherical_harmonics.jl:9-31)."""

    def __init__(self, coeff, LMAX, KMAX, rmax):
        coeff = np.ascontiguousarray(coeff, dtype=np.complex128)
        if coeff.shape != (KMAX, LMAX * LMAX):
            raise ValueError(f"coeff must be (KMAX, LMAX^2) = ({KMAX}, {LMAX * LMAX}), got {coeff.shape}")
        self.coeff, self.LMAX, self.KMAX, self.rmax = coeff, LMAX, KMAX, float(rmax)

    def dq(self):
        """spherical_harmonics.jl:102"""
        return self.rmax / self.KMAX

    def copy(self):
        return SphericalHarmonicsVolume(self.coeff.copy(), self.LMAX, self.KMAX, self.rmax)


class HipContext:
    """Device context (one HIP stream).  HIP_init() returns the default one."""

    def __init__(self, device=0):
        h = C.c_void_p()
        check(lib().tp_init(device, C.byref(h)), "tp_init")
        self.h = h
        self.device = device
        track(self, 2)

    def close(self):
        if getattr(self, "h", None):
            lib().tp_finalize(self.h)
            self.h = None

    _release = close

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_timing(self, on=True):
        check(lib().tp_set_timing(self.h, int(on)), "tp_set_timing")

    def last_timings(self):
        out = (C.c_double * 4)()
        check(lib().tp_last_timings(self.h, out, 4), "tp_last_timings")
        return list(out)


_default_ctx = None


def HIP_init(device=0):
    """~ CUDA_init (cuda.jl:10-21): open device 0 and keep it as the default context."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = HipContext(device)
    return _default_ctx


def _ctx(ctx):
    return ctx if ctx is not None else HIP_init()


class BasisTypeHip:
    """~ BasisTypeCuda (cuda.jl:24-42): the precomputed basis, resident on the GPU."""

    def __init__(self, L, LMAX, N, K, lam=0.0, dq=0.0, forIntensity=True, ctx=None, _host=None):
        self.ctx = _ctx(ctx)
        h = C.c_void_p()
        if _host is None:
            check(lib().tp_basis_create(self.ctx.h, L, LMAX, N, K, float(lam), float(dq), int(bool(forIntensity)),
                                        C.byref(h)), "tp_basis_create")
        else:
            w, ind, pac = _host
            pp = pac.ctypes.data if pac is not None else None
            npc = pac.shape[0] if pac is not None else 0
            check(lib().tp_basis_from_host(self.ctx.h, w.ctypes.data, ind.ctypes.data, w.shape[0], pp, npc, N, L,
                                           LMAX, K, float(lam), float(dq), C.byref(h)), "tp_basis_from_host")
        self.h = h
        track(self, 1)
        self.L, self.LMAX, self.N, self.K, self.lam, self.dq = L, LMAX, N, K, float(lam), float(dq)
        self.lrange = range(0, L + 1, 2 if forIntensity else 1)
        bl, nl, kl, R = C.c_int64(), C.c_int64(), C.c_int64(), C.c_int64()
        check(lib().tp_basis_info(h, C.byref(bl), C.byref(nl), C.byref(kl), C.byref(R)), "tp_basis_info")
        self.basislen, self.n_lt, self.klength, self.R = bl.value, nl.value, kl.value, R.value
        self._refs = {}

    def _release(self):
        for r in getattr(self, "_refs", {}).values():
            lib().tp_ref_destroy(r[0])
        self._refs = {}
        if getattr(self, "h", None):
            lib().tp_basis_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def set_precision(self, mode):
        """'fp64' (default), 'fp16x3' (compensated FP16 MFMA triple products and
        contraction), 'bf16', 'triple16' (compensated FP16 triple products,
        FP64 contraction) or 'fp16x3_t64' (triple products on the sorted
        l-triples, FP32 MFMA stage 1 and compensated FP16 MFMA stage 2,
        G = H.As on compensated FP16 MFMA, the FP64 separated-grid
        contraction c3 = Bm.G and energy).  On the curved Ewald sphere
        (lambda dq != 0) the modes are 'fp64' and 'fp16x3_t64', whose G is
        then formed in FP64 (k_glam)."""
        m = {"fp64": PREC_FP64, "fp16x3": PREC_FP16X3, "bf16": PREC_BF16,
             "triple16": 3, "fp16x3_t64": 4}[mode] if isinstance(mode, str) else mode
        check(lib().tp_basis_set_precision(self.h, m), "tp_basis_set_precision")

    def precision(self):
        """The current precision mode (PREC_* code)."""
        return int(lib().tp_basis_get_precision(self.h))

    def histogram(self, c3ref, K3_range=None):
        """A device copy of the normalised histogram slices visited by energy()'s
        walk (tp_ref_create), as a HistogramRef: pass it where a c3ref is
        expected to skip any host work per call.  It holds the values of
        c3ref at this call; later changes to the array are not seen by it."""
        return HistogramRef(self, self.ref(c3ref, K3_range), K3_range)

    def ref(self, c3ref, K3_range=None):
        """The tp_ref handle for c3ref (a HistogramRef of this basis, or an
        array).  Arrays are cached by their contents (a 64-bit hash of their
        bytes), so an array changed in place gets a fresh device copy."""
        if isinstance(c3ref, HistogramRef):
            if c3ref.basis is not self:
                raise ValueError("HistogramRef belongs to another basis")
            if K3_range is not None and K3_range != c3ref.K3_range:
                raise ValueError("HistogramRef was made for another K3_range")
            return c3ref.h
        if K3_range is None:
            K3_range = range(1, self.K + 1)
        arr = np.asarray(c3ref, dtype=np.float64)
        if arr.ndim != 5 or arr.shape[0] != self.N or arr.shape[1] != 2 * self.N:
            raise ValueError("c3ref must be (N, 2N, Kr, Kr, Kr)")
        if arr.flags.c_contiguous or arr.flags.f_contiguous:
            key = (_digest(arr.ravel(order="K")), arr.shape, arr.flags.c_contiguous, K3_range.start, K3_range.stop)
        else:
            key = (_digest(np.ascontiguousarray(arr)), arr.shape, True, K3_range.start, K3_range.stop)
        hit = self._refs.get(key)
        if hit is not None:
            return hit[0]
        flat = np.ascontiguousarray(arr.ravel(order="F"))
        Kr = arr.shape[2]
        h = C.c_void_p()
        check(lib().tp_ref_create(self.ctx.h, self.h, flat.ctypes.data, Kr, K3_range.start, K3_range.stop - 1,
                                  C.byref(h)), "tp_ref_create")
        self._refs[key] = (h, None)
        return h


class HistogramRef:
    """A histogram resident on the device for one basis and K3_range (the
    compressed c3ref of new_energy, cuda.jl:127-139).  Kept alive by the
    basis; pass it as c3ref to energy(), DeviceState.energy() and the
    annealing drivers."""

    def __init__(self, basis, h, K3_range):
        self.basis, self.h = basis, h
        self.K3_range = K3_range if K3_range is not None else range(1, basis.K + 1)


def HIP_calculate_basis(L, LMAX, N, K, lam=0.0, dq=0.0, forIntensity=True, ctx=None):
    """~ CUDA_calculate_basis (cuda.jl:57-65)."""
    return BasisTypeHip(L, LMAX, N, K, lam, dq, forIntensity, ctx)


def HIP_store_basis(wignerlist, indices, N, L, LMAX, K, pacombos=None, lam=0.0, dq=0.0, ctx=None):
    """~ CUDA_store_basis (cuda.jl:45-55): a basis from the caller's wignerlist
    and indices (basislen x 9 rows [sean1, sean2, sean3, l1, m1, l2, m2, l3, -m3])."""
    w = np.ascontiguousarray(wignerlist, dtype=np.float64)
    ind = np.ascontiguousarray(indices, dtype=np.int64).reshape(w.shape[0], 9)
    pac = None if pacombos is None else np.ascontiguousarray(pacombos, dtype=np.int64).reshape(-1, 9)
    return BasisTypeHip(L, LMAX, N, K, lam, dq, True, ctx, _host=(w, ind, pac))


def _coeff_ptr(intensity, basis, nshell=None):
    c = intensity.coeff if isinstance(intensity, SphericalHarmonicsVolume) else np.asarray(intensity)
    c = np.ascontiguousarray(c, dtype=np.complex128)
    need = basis.K if nshell is None else nshell
    if c.ndim != 2 or c.shape[0] < need or c.shape[1] < basis.LMAX * basis.LMAX:
        raise ValueError(f"coefficients must be (>= {need}, >= LMAX^2) complex")
    return c, c.shape[1]


def FullCorrelation_parallized(intensity, basis, minimal=True, normalize=False, return_raw=False):
    """c3 on the GPU (cuda.jl:74-108).  return_raw gives the (R, klength) matrix."""
    c, ld = _coeff_ptr(intensity, basis)
    res = np.zeros((basis.klength, basis.R))      # column kt contiguous == Julia R x klength
    check(lib().tp_c3(basis.ctx.h, basis.h, c.ctypes.data, ld, res.ctypes.data), "tp_c3")
    res = res.T
    if return_raw:
        return res
    N, K = basis.N, basis.K
    t = np.zeros((N, 2 * N, K, K, K))
    i = 0
    for k1 in range(1, K + 1):
        for k2 in range(1, k1 + 1):
            for k3 in range(1, k2 + 1):
                t[:, :, k3 - 1, k2 - 1, k1 - 1] = res[:, i].reshape(2 * N, N).T
                i += 1
    return t / np.abs(t).sum() if normalize else t


def energy(intensity, basis, c3ref, K3_range=None, measure="Bayes", negativity_factor=0.0):
    """Bayes energy (correlations.jl:196-226).  negativity_factor is ignored, as in the reference."""
    ref = basis.ref(c3ref, K3_range)
    c, ld = _coeff_ptr(intensity, basis)
    E = C.c_double()
    check(lib().tp_energy(basis.ctx.h, basis.h, ref, c.ctypes.data, ld, 0 if measure == "Bayes" else 1,
                          C.byref(E)), "tp_energy")
    return E.value


class DeviceState:
    """Device-resident coefficients of all KMAX shells (the annealing state)."""

    def __init__(self, intensity, basis, nshell=None):
        nshell = intensity.KMAX if (nshell is None and isinstance(intensity, SphericalHarmonicsVolume)) else nshell
        c, ld = _coeff_ptr(intensity, basis, nshell)
        nshell = c.shape[0] if nshell is None else nshell
        h = C.c_void_p()
        check(lib().tp_state_create(basis.ctx.h, basis.h, c.ctypes.data, ld, nshell, C.byref(h)), "tp_state_create")
        self.h, self.basis, self.nshell, self.ld = h, basis, nshell, ld
        track(self, 0)
        self.template = c[:nshell].copy()

    def _release(self):
        if getattr(self, "h", None):
            lib().tp_state_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def read(self):
        out = self.template.copy()
        check(lib().tp_state_read(self.basis.ctx.h, self.h, out.ctypes.data, self.ld), "tp_state_read")
        return out

    def energy(self, c3ref, K3_range=None):
        ref = self.basis.ref(c3ref, K3_range)
        E = C.c_double()
        check(lib().tp_state_energy(self.basis.ctx.h, self.basis.h, ref, self.h, C.byref(E)), "tp_state_energy")
        return E.value

    def c3(self):
        res = np.zeros((self.basis.klength, self.basis.R))
        check(lib().tp_state_c3(self.basis.ctx.h, self.basis.h, self.h, res.ctypes.data), "tp_state_c3")
        return res.T


def rotate_all_at_once(state, basis, c3ref, nsteps, initial_stepsize=np.pi / 4, stepsizefactor=1.01,
                       initial_temperature_factor=0.1, temperature_decay=0.99998, seed=0, ntemp=50,
                       K3_range=None):
    """Plain MC simulated annealing on the device (determination.jl:281-338)."""
    ref = basis.ref(c3ref, K3_range)
    p = AnnealParams(nsteps, ntemp, initial_stepsize, stepsizefactor, initial_temperature_factor,
                     temperature_decay, seed)
    E = np.zeros(max(nsteps, 1))
    En = np.zeros(max(nsteps, 1))
    acc = np.zeros(max(nsteps, 1), np.int32)
    sc = np.zeros(4)
    check(lib().tp_anneal(basis.ctx.h, basis.h, ref, state.h, C.byref(p), E.ctypes.data, En.ctypes.data,
                          acc.ctypes.data, sc.ctypes.data), "tp_anneal")
    return dict(E=E[:nsteps], E_new=En[:nsteps], accepted=acc[:nsteps], E0=sc[0], T0=sc[1], E_final=sc[2],
                T_final=sc[3])


def intensity_from_atoms(atoms, KMAX, qmax, LMAX, lfill, ntheta, nphi, ctx=None):
    """SH coefficients of point-atom intensity on shells q_k = k*dq, dq = qmax/KMAX.

    As in createSphericalHarmonicsStructure (structure.jl:74-134) the returned
    volume stores qmax in its rmax field, so dq(volume) = qmax/KMAX
    (spherical_harmonics.jl:102-104)."""
    ctx = _ctx(ctx)
    atoms = np.ascontiguousarray(atoms, dtype=np.float64)
    out = np.zeros((KMAX, LMAX * LMAX), np.complex128)
    dq = qmax / KMAX
    check(lib().tp_intensity_from_atoms(ctx.h, atoms.ctypes.data, atoms.shape[0], KMAX, dq, LMAX, lfill, ntheta,
                                        nphi, out.ctypes.data), "tp_intensity_from_atoms")
    return SphericalHarmonicsVolume(out, LMAX, KMAX, qmax)


def rotate(intensity, basis, quats, nshell=None):
    """Rotate the even l >= 2 blocks of every shell by one unit quaternion per l
    (quats shape (L/2, 4), w,x,y,z; a single (4,) quaternion is used for every l),
    in FP64 through the real Wigner-D recurrence.  Returns the rotated (nshell, ld)
    coefficient array; f'(r) = f(R^-1 r)."""
    c, ld = _coeff_ptr(intensity, basis, nshell)
    nshell = c.shape[0] if nshell is None else nshell
    q = np.asarray(quats, dtype=np.float64)
    if q.ndim == 1:
        q = np.tile(q, (basis.L // 2, 1))
    q = np.ascontiguousarray(q.reshape(basis.L // 2, 4))
    out = np.zeros_like(c[:nshell])
    check(lib().tp_rotate(basis.ctx.h, basis.h, c.ctypes.data, ld, nshell, q.ctypes.data, out.ctypes.data),
          "tp_rotate")
    return out


def energy_grad(intensity, basis, c3ref, K3_range=None):
    """(E, gamma): the FP64 Bayes energy (as energy()) and its gradient
    gamma[k, s] = dE/dRe c_k[s] + i dE/dIm c_k[s] for the first K shells
    (tp_energy_grad; every coefficient entry an independent complex variable)."""
    c, ld = _coeff_ptr(intensity, basis)
    ref = basis.ref(c3ref, K3_range)
    E = C.c_double()
    g = np.zeros((basis.K, ld), np.complex128)
    check(lib().tp_energy_grad(basis.ctx.h, basis.h, ref.value, c.ctypes.data, ld, C.byref(E), g.ctypes.data),
          "tp_energy_grad")
    return E.value, g


class PhotonHistogram:
    """Photon-pair / photon-triplet histograms of sparse images on the device
    (tp_hist_*): add() is histogramCorrelationsInPicture_alltoall
    (datagen.jl:228-282) over many pictures; read() gives c1, c2, c3 with the
    reference's index order (c2[ai, k2, k1], c3[ai, bi, k3, k2, k1], zero-based)
    and the integer counts behind them."""

    def __init__(self, N, K2, K3, dq, ctx=None):
        self.ctx = ctx or HIP_init()
        self.N, self.K2, self.K3, self.dq = N, K2, K3, float(dq)
        h = C.c_void_p()
        check(lib().tp_hist_create(self.ctx.h, N, K2, K3, self.dq, C.byref(h)), "tp_hist_create")
        self.h = h
        track(self, 0)

    def _release(self):
        if getattr(self, "h", None):
            lib().tp_hist_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def add(self, photons, offsets):
        """photons (n, 3); picture i = photons[offsets[i]:offsets[i+1]]."""
        ph = np.ascontiguousarray(photons, dtype=np.float64).reshape(-1, 3)
        off = np.ascontiguousarray(offsets, dtype=np.int64)
        check(lib().tp_hist_add(self.h, ph.ctypes.data, off.ctypes.data, len(off) - 1), "tp_hist_add")

    def read(self):
        N, K2, K3 = self.N, self.K2, self.K3
        c1, c2, c3 = np.zeros(K2), np.zeros(N * K2 * K2), np.zeros(2 * N * N * K3 ** 3)
        n1 = np.zeros(K2, np.uint64)
        n2 = np.zeros(N * K2 * K2, np.uint64)
        n3 = np.zeros(2 * N * N * K3 ** 3, np.uint64)
        check(lib().tp_hist_read(self.h, c1.ctypes.data, c2.ctypes.data, c3.ctypes.data, n1.ctypes.data,
                                 n2.ctypes.data, n3.ctypes.data), "tp_hist_read")
        s2, s3 = (N, K2, K2), (N, 2 * N, K3, K3, K3)
        return {"c1": c1, "c2": c2.reshape(s2, order="F"), "c3": c3.reshape(s3, order="F"), "n1": n1,
                "n2": n2.reshape(s2, order="F"), "n3": n3.reshape(s3, order="F")}


def twoPhotons(volume, basis, K2, minimal=False, normalize=False, dq=None):
    """Two-photon correlation c2 (N, K2, K2), indexed [a, k2, k1] (zero-based),
    of shells 1..K2 of `volume` on the GPU (correlations.jl:229-243).  The l
    range, N and lambda are the basis's; with lambda != 0 the series is taken
    at alpha_star with the volume's shell spacing (dq(volume), or `dq` for a
    plain coefficient array; tp_two_photons_ewald), else tp_two_photons."""
    c, ld = _coeff_ptr(volume, basis, K2)
    out = np.zeros(basis.N * K2 * K2)
    if basis.lam != 0.0:
        if dq is None:
            if not isinstance(volume, SphericalHarmonicsVolume):
                raise ValueError("lambda != 0 needs the volume's shell spacing: pass dq")
            dq = volume.dq()
        check(lib().tp_two_photons_ewald(basis.ctx.h, basis.h, c.ctypes.data, ld, K2, int(bool(minimal)),
                                         int(bool(normalize)), float(dq), out.ctypes.data), "tp_two_photons_ewald")
    else:
        check(lib().tp_two_photons(basis.ctx.h, basis.h, c.ctypes.data, ld, K2, int(bool(minimal)),
                                   int(bool(normalize)), out.ctypes.data), "tp_two_photons")
    return out.reshape((basis.N, K2, K2), order="F")


def retrieveSolution(c2, L, LMAX, K2_range, qmax, lam=0.0, return_info=False):
    """A start structure from a two-photon correlation (tp_retrieve_solution;
    correlations.jl:262-319): a SphericalHarmonicsVolume of max(K2_range)
    shells whose c2 matches c2 ([a, k2, k1]) on K2_range.  K2_range is an int
    (1..K2) or a range of 1-based shells; lam != 0 fits at alpha_star
    (tp_retrieve_solution_ewald)."""
    rng = range(1, K2_range + 1) if isinstance(K2_range, int) else K2_range
    lo, hi = min(rng), max(rng)
    c2 = np.asarray(c2, dtype=np.float64)
    N, K2a, K2b = c2.shape
    if K2a != K2b:
        raise ValueError("c2 must be (N, K2, K2)")
    flat = np.ascontiguousarray(c2.ravel(order="F"))
    out = np.zeros((hi, LMAX * LMAX), np.complex128)
    flipped, neg = C.c_int(), C.c_double()
    if lam != 0.0:
        check(lib().tp_retrieve_solution_ewald(flat.ctypes.data, N, K2a, lo, hi, L, LMAX, float(qmax), float(lam),
                                               out.ctypes.data, LMAX * LMAX, C.byref(flipped), C.byref(neg)),
              "tp_retrieve_solution_ewald")
    else:
        check(lib().tp_retrieve_solution(flat.ctypes.data, N, K2a, lo, hi, L, LMAX, out.ctypes.data, LMAX * LMAX,
                                         C.byref(flipped), C.byref(neg)), "tp_retrieve_solution")
    vol = SphericalHarmonicsVolume(out, LMAX, hi, qmax)
    if return_info:
        return vol, {"flipped": bool(flipped.value), "negativity": neg.value}
    return vol


def energy_batch(intensities, bases, c3refs, K3_ranges=None):
    """Bayes energies of many evaluations in one call (tp_energy_batch): one per
    (intensity, basis, c3ref); bases share L, LMAX, N and precision, K may differ."""
    n = len(intensities)
    K3_ranges = K3_ranges or [None] * n
    refs = (C.c_void_p * n)(*[bases[i].ref(c3refs[i], K3_ranges[i]).value for i in range(n)])
    hb = (C.c_void_p * n)(*[b.h.value for b in bases])
    arrs = [_coeff_ptr(intensities[i], bases[i]) for i in range(n)]
    ptrs = (C.c_void_p * n)(*[a[0].ctypes.data for a in arrs])
    lds = (C.c_int64 * n)(*[a[1] for a in arrs])
    E = np.zeros(n)
    check(lib().tp_energy_batch(bases[0].ctx.h, n, hb, refs, ptrs, lds, E.ctypes.data), "tp_energy_batch")
    return E


class AnnealBatch:
    """Concurrent annealing runs scored together (tp_batch_*): run g has its own
    histogram c3refs[g] (shape (N, 2N, K_g, K_g, K_g); shell count K_g), start
    coefficients starts[g] and seed; every sweep scores nprop rotation proposals
    per run in one graph replay.  `basis` (mixed precision, K >= every K_g) holds
    the tables; the histograms are walked over K3_range 1:K_g on it."""

    def __init__(self, basis, c3refs, starts, seeds, nprop=64, hist_cap=0):
        self.basis, self.nprop = basis, nprop
        self.ngroups = n = len(c3refs)
        self.Ks = [(c3refs[g].K3_range.stop - 1) if isinstance(c3refs[g], HistogramRef)
                   else np.asarray(c3refs[g]).shape[2] for g in range(n)]
        refs = (C.c_void_p * n)(*[basis.ref(c3refs[g], range(1, self.Ks[g] + 1)).value for g in range(n)])
        arrs = [_coeff_ptr(starts[g], basis, self.Ks[g]) for g in range(n)]
        self._starts = [a[0][:self.Ks[g]].copy() for g, a in enumerate(arrs)]
        ptrs = (C.c_void_p * n)(*[a[0].ctypes.data for a in arrs])
        lds = (C.c_int64 * n)(*[a[1] for a in arrs])
        self._lds = [a[1] for a in arrs]
        sd = (C.c_uint64 * n)(*[int(s) for s in seeds])
        h = C.c_void_p()
        check(lib().tp_batch_create(basis.ctx.h, basis.h, n, nprop, refs, ptrs, lds, sd, hist_cap, C.byref(h)),
              "tp_batch_create")
        self.h, self.hist_cap = h, hist_cap
        track(self, 0)

    def _release(self):
        if getattr(self, "h", None):
            lib().tp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def autotune(self, reps=3):
        ms = np.zeros(14)
        cfg = np.zeros(4, np.int32)
        check(lib().tp_batch_autotune(self.h, reps, ms.ctypes.data, cfg.ctypes.data), "tp_batch_autotune")
        if int(cfg[3]) >= 4:
            # factored contraction (fp16x3_t64): k_hgemm_g candidates (all row
            # blocks x 64 or 32 columns, 3 row blocks x 64 columns, 4 row blocks
            # x 16 columns per workgroup), each followed by k_c3sep
            names = ["allrows64/factored", "allrows32/factored", "rows3x64/factored", "rows4x16/factored"]
            tile_ms = {n: float(v) for n, v in zip(names, ms[2:6])}
            return dict(order_ms=ms[:2].tolist(), tile_ms=tile_ms, tile=(int(cfg[0]), int(cfg[1])),
                        order=int(cfg[2]), schedule="factored", candidate=names[int(cfg[3]) - 4])
        names = [f"{t}/{s}" for s in ("rows", "walk", "xcdrows") for t in ("128x64", "64x64", "128x128", "64x128")]
        return dict(order_ms=ms[:2].tolist(), tile_ms=dict(zip(names, ms[2:].tolist())),
                    tile=(int(cfg[0]), int(cfg[1])), order=int(cfg[2]),
                    schedule={1: "rows", 2: "walk", 3: "xcdrows"}[int(cfg[3])],
                    candidate="%dx%d/%s" % (int(cfg[0]), int(cfg[1]), {1: "rows", 2: "walk", 3: "xcdrows"}[int(cfg[3])]))

    def set_contraction(self, cand):
        """Fix the contraction candidate (0..11 in autotune()'s tile_ms order; 0..3 for fp16x3_t64)."""
        check(lib().tp_batch_set_contraction(self.h, int(cand)), "tp_batch_set_contraction")

    def probe(self, temperature_factor=0.1):
        check(lib().tp_batch_probe(self.h, float(temperature_factor)), "tp_batch_probe")

    def run(self, nsweeps, decay=0.99998, factor=1.01):
        check(lib().tp_batch_run(self.h, int(nsweeps), float(decay), float(factor)), "tp_batch_run")

    def group(self, g):
        E, T, a = C.c_double(), C.c_double(), C.c_double()
        na, ns = C.c_int32(), C.c_int32()
        check(lib().tp_batch_group(self.h, g, C.byref(E), C.byref(T), C.byref(a), C.byref(na), C.byref(ns)),
              "tp_batch_group")
        return dict(E=E.value, T=T.value, amax=a.value, nacc=na.value, nsweep=ns.value)

    def coeff(self, g):
        out = self._starts[g].copy()
        check(lib().tp_batch_coeff(self.h, g, out.ctypes.data, self._lds[g]), "tp_batch_coeff")
        return out

    def history(self, n=None):
        n = self.hist_cap if n is None else min(n, self.hist_cap)
        out = np.zeros((max(n, 0), self.ngroups))
        if n > 0:
            check(lib().tp_batch_history(self.h, out.ctypes.data, n), "tp_batch_history")
        return out

    def eval_quats(self, quats):
        q = np.ascontiguousarray(np.asarray(quats, dtype=np.float64).reshape(self.ngroups * self.nprop,
                                                                             self.basis.L // 2, 4))
        E = np.zeros(self.ngroups * self.nprop)
        check(lib().tp_batch_eval_quats(self.h, q.ctypes.data, E.ctypes.data), "tp_batch_eval_quats")
        return E

    def eval_rotations(self, ngeo, gen=None, theta=None, uvb=None):
        """Energies of given rotations of the current states (tp_batch_eval_rotations):
        slots p < ngeo turn block l by exp(-theta G_l) with gen (ngroups,
        sum_l (2l+1)^2) run g's real skew-symmetric generators (row-major,
        concatenated over l = 2..L, Frobenius norm <= 1) and theta
        (ngroups*nprop, L/2); the other slots by the plane rotation of uvb
        (ngroups*nprop, L/2, 2(2L+1)+1) = (u, v, b).  All in the basis
        x = Umat(l) c."""
        n, L = self.ngroups * self.nprop, self.basis.L
        gsz = sum((2 * l + 1) ** 2 for l in range(2, L + 1, 2))
        E = np.zeros(n)
        args = [None if a is None else np.ascontiguousarray(a, dtype=np.float64) for a in (gen, theta, uvb)]
        if ngeo > 0 and (args[0] is None or args[0].size != self.ngroups * gsz):
            raise ValueError("gen must hold ngroups x sum_l (2l+1)^2 doubles")
        if args[1] is not None and args[1].shape != (n, L // 2):
            raise ValueError("theta must be (ngroups*nprop, L/2)")
        if args[2] is not None and args[2].shape != (n, L // 2, 2 * (2 * L + 1) + 1):
            raise ValueError("uvb must be (ngroups*nprop, L/2, 2(2L+1)+1)")
        ptr = [None if a is None else a.ctypes.data for a in args]
        check(lib().tp_batch_eval_rotations(self.h, int(ngeo), *ptr, E.ctypes.data), "tp_batch_eval_rotations")
        return E

    def set_state(self, g, coeff, E):
        """Run g's current structure and energy (tp_batch_set_state)."""
        c = np.ascontiguousarray(coeff, dtype=np.complex128)
        check(lib().tp_batch_set_state(self.h, int(g), c.ctypes.data, c.shape[1], float(E)), "tp_batch_set_state")


def _ab_log(self, capacity):
    check(lib().tp_batch_log(self.h, int(capacity)), "tp_batch_log")


def _ab_log_count(self):
    n = C.c_int32()
    check(lib().tp_batch_log_count(self.h, C.byref(n)), "tp_batch_log_count")
    return n.value


def _ab_log_read(self, first, n):
    X = np.zeros((n, 192), np.float32)
    T = np.zeros((n, 36), np.float32)
    check(lib().tp_batch_log_read(self.h, first, n, X.ctypes.data, T.ctypes.data), "tp_batch_log_read")
    return X, T


def _ab_set_proposer(self, mlp, nlearn, noise=0.25):
    check(lib().tp_batch_set_proposer(self.h, mlp.h if mlp is not None else None, int(nlearn), float(noise)),
          "tp_batch_set_proposer")
    self._mlp = mlp


def _ab_set_amax(self, amax):
    check(lib().tp_batch_set_amax(self.h, float(amax)), "tp_batch_set_amax")


def _ab_set_family(self, family):
    """'plane' (the reference's move, default) or 'quat' (SO(3) per l)."""
    f = {"quat": 0, "plane": 1}[family] if isinstance(family, str) else int(family)
    check(lib().tp_batch_set_family(self.h, f), "tp_batch_set_family")


AnnealBatch.set_family = _ab_set_family
AnnealBatch.set_amax = _ab_set_amax
AnnealBatch.log = _ab_log
AnnealBatch.log_count = _ab_log_count
AnnealBatch.log_read = _ab_log_read
AnnealBatch.set_proposer = _ab_set_proposer


class MlpProposer:
    """The learned rotation proposer (tp_mlp_*): BF16 MLP 192 -> 256x3 -> 64
    whose first 36 outputs are one quaternion per even l = 2..18."""

    def __init__(self, seed=0, ctx=None):
        self.ctx = _ctx(ctx)
        h = C.c_void_p()
        check(lib().tp_mlp_create(self.ctx.h, int(seed), C.byref(h)), "tp_mlp_create")
        self.h = h
        track(self, 0)
        self.nparam = lib().tp_mlp_nparam()

    def _release(self):
        if getattr(self, "h", None):
            lib().tp_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass

    def params(self):
        p = np.zeros(self.nparam, np.float32)
        check(lib().tp_mlp_get_params(self.h, p.ctypes.data), "tp_mlp_get_params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, dtype=np.float32)
        check(lib().tp_mlp_set_params(self.h, p.ctypes.data), "tp_mlp_set_params")

    def forward(self, X):
        X = np.ascontiguousarray(X, dtype=np.float32)
        Y = np.zeros((X.shape[0], 64), np.float32)
        check(lib().tp_mlp_forward(self.h, X.ctypes.data, X.shape[0], Y.ctypes.data), "tp_mlp_forward")
        return Y

    def loss_grad(self, X, T):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        g = np.zeros(self.nparam, np.float32)
        loss = C.c_double()
        check(lib().tp_mlp_loss_grad(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], C.byref(loss), g.ctypes.data),
              "tp_mlp_loss_grad")
        return loss.value, g

    def train(self, X, T, steps, batch=512, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, dtype=np.float32)
        T = np.ascontiguousarray(T, dtype=np.float32)
        tr = np.zeros(max(steps, 1), np.float32)
        check(lib().tp_mlp_train(self.h, X.ctypes.data, T.ctypes.data, X.shape[0], steps, batch, lr, seed,
                                 tr.ctypes.data), "tp_mlp_train")
        return tr[:steps]

    def train_log(self, anneal_batch, steps, batch=512, lr=1e-3, seed=0):
        tr = np.zeros(max(steps, 1), np.float32)
        check(lib().tp_mlp_train_log(self.h, anneal_batch.h, steps, batch, lr, seed, tr.ctypes.data),
              "tp_mlp_train_log")
        return tr[:steps]
