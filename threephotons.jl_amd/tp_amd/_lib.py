"""Build and load libtp.so (the gfx950 C-ABI library declared in include/tp.h)."""
import atexit
import ctypes as C
import os
import subprocess
import weakref

PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))   # threephotons.jl_amd/
REPO_ROOT = os.path.dirname(PKG_ROOT)
LIB_PATH = os.path.join(PKG_ROOT, "libtp.so")
SRC = os.path.join(PKG_ROOT, "csrc", "tp_hip.hip")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    "-Wno-unused-value", "-Wno-unused-result",
    # packed FP32 instructions are kept out of every kernel
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def build(verbose=False):
    """Compile csrc/tp_hip.hip for gfx950 into threephotons.jl_amd/libtp.so."""
    tmp = LIB_PATH + ".tmp"
    cmd = [HIPCC] + HIPCC_FLAGS + ["-o", tmp, SRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + r.stderr[-4000:])
    os.replace(tmp, LIB_PATH)
    if verbose:
        print("built", LIB_PATH)
    return LIB_PATH


_lib = None

# Handles still alive when the interpreter exits are released here, before
# module teardown clears the names their finalizers use (run objects first,
# then bases, then contexts), so no device object outlives the process's
# HIP runtime.
_live = ([], [], [])


def track(obj, tier):
    """Release obj (its _release()) at exit in tier order: 0 runs, batches,
    networks, states; 1 bases; 2 contexts."""
    lst = _live[tier]
    if len(lst) > 4096:
        lst[:] = [w for w in lst if w() is not None]
    lst.append(weakref.ref(obj))


def _release_all():
    for lst in _live:
        for w in lst:
            o = w()
            if o is not None:
                try:
                    o._release()
                except Exception:
                    pass
        lst.clear()


atexit.register(_release_all)


class TPError(RuntimeError):
    pass


def lib():
    """The loaded libtp.so; raises if it has not been built (no fallback)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise TPError(f"libtp.so not built ({LIB_PATH}); run __graft_entry__.build()")
    L = C.CDLL(LIB_PATH)
    vp, i32, i64, dbl = C.c_void_p, C.c_int, C.c_int64, C.c_double
    pp = C.POINTER(C.c_void_p)
    L.tp_abi_version.restype = i32
    L.tp_last_error.restype = C.c_char_p
    L.tp_init.argtypes = [i32, pp]
    L.tp_finalize.argtypes = [vp]
    L.tp_finalize.restype = None
    L.tp_basis_create.argtypes = [vp, i32, i32, i32, i32, dbl, dbl, i32, pp]
    L.tp_basis_from_host.argtypes = [vp, vp, vp, i64, vp, i64, i32, i32, i32, i32, dbl, dbl, pp]
    L.tp_basis_destroy.argtypes = [vp]
    L.tp_basis_destroy.restype = None
    L.tp_basis_info.argtypes = [vp, C.POINTER(i64), C.POINTER(i64), C.POINTER(i64), C.POINTER(i64)]
    L.tp_c3.argtypes = [vp, vp, vp, i64, vp]
    L.tp_ref_create.argtypes = [vp, vp, vp, i32, i32, i32, pp]
    L.tp_ref_destroy.argtypes = [vp]
    L.tp_ref_destroy.restype = None
    L.tp_energy.argtypes = [vp, vp, vp, vp, i64, i32, C.POINTER(dbl)]
    L.tp_state_create.argtypes = [vp, vp, vp, i64, i32, pp]
    L.tp_state_destroy.argtypes = [vp]
    L.tp_state_destroy.restype = None
    L.tp_state_read.argtypes = [vp, vp, vp, i64]
    L.tp_state_energy.argtypes = [vp, vp, vp, vp, C.POINTER(dbl)]
    L.tp_state_c3.argtypes = [vp, vp, vp, vp]
    L.tp_anneal.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.tp_intensity_from_atoms.argtypes = [vp, vp, i32, i32, dbl, i32, i32, i32, i32, vp]
    L.tp_basis_set_precision.argtypes = [vp, i32]
    L.tp_basis_get_precision.argtypes = [vp]
    L.tp_set_timing.argtypes = [vp, i32]
    L.tp_last_timings.argtypes = [vp, C.POINTER(dbl), i32]
    L.tp_energy_batch.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.tp_energy_grad.argtypes = [vp, vp, vp, vp, i64, vp, vp]
    L.tp_hist_create.argtypes = [vp, i32, i32, i32, dbl, pp]
    L.tp_hist_add.argtypes = [vp, vp, vp, i32]
    L.tp_hist_read.argtypes = [vp, vp, vp, vp, vp, vp, vp]
    L.tp_hist_destroy.argtypes = [vp]
    L.tp_hist_destroy.restype = None
    L.tp_two_photons.argtypes = [vp, vp, vp, i64, i32, i32, i32, vp]
    L.tp_retrieve_solution.argtypes = [vp, i32, i32, i32, i32, i32, i32, vp, i64, vp, vp]
    L.tp_two_photons_ewald.argtypes = [vp, vp, vp, i64, i32, i32, i32, dbl, vp]
    L.tp_retrieve_solution_ewald.argtypes = [vp, i32, i32, i32, i32, i32, i32, dbl, dbl, vp, i64, vp, vp]
    L.tp_fit_create.argtypes = [vp, vp, vp, i64, i32, i32, pp]
    L.tp_fit_set_volume.argtypes = [vp, vp, i64]
    L.tp_fit_similarity.argtypes = [vp, i32, vp, vp]
    L.tp_fit_destroy.argtypes = [vp]
    L.tp_fit_destroy.restype = None
    L.tp_sh_surface.argtypes = [vp, i32, i32, vp, i64, vp]
    L.tp_sh_coeff.argtypes = [vp, i32, i32, vp, vp, i64]
    L.tp_sph_bessel.argtypes = [vp, i32, i32, i32, vp, i64, vp]
    L.tp_density_surface.argtypes = [vp, vp, i32, i32, i32, dbl, vp]
    L.tp_absolute_square.argtypes = [vp, i32, i32, vp, i64, vp]
    L.tp_structure_from_atoms.argtypes = [vp, vp, i32, i32, i32, dbl, vp, vp, vp, i64]
    L.tp_sraar.argtypes = [vp, i32, i32, vp, i64, i32, dbl, dbl, dbl, dbl, vp, C.c_uint64, vp, vp]
    L.tp_pictures.argtypes = [vp, i32, i32, dbl, vp, i64, vp, i32, i32, vp, vp, vp]
    L.tp_hist_add_generated.argtypes = [vp, i32, i32, dbl, vp, i64, vp, i32, i32, vp]
    L.tp_batch_create.argtypes = [vp, vp, i32, i32, vp, vp, vp, vp, i32, pp]
    L.tp_batch_destroy.argtypes = [vp]
    L.tp_batch_destroy.restype = None
    L.tp_batch_autotune.argtypes = [vp, i32, vp, vp]
    L.tp_batch_probe.argtypes = [vp, dbl]
    L.tp_batch_run.argtypes = [vp, i32, dbl, dbl]
    L.tp_batch_group.argtypes = [vp, i32, vp, vp, vp, vp, vp]
    L.tp_batch_coeff.argtypes = [vp, i32, vp, i64]
    L.tp_batch_history.argtypes = [vp, vp, i32]
    L.tp_batch_eval_quats.argtypes = [vp, vp, vp]
    L.tp_rotate.argtypes = [vp, vp, vp, i64, i32, vp, vp]
    L.tp_mlp_nparam.restype = i32
    L.tp_mlp_create.argtypes = [vp, C.c_uint64, pp]
    L.tp_mlp_destroy.argtypes = [vp]
    L.tp_mlp_destroy.restype = None
    L.tp_mlp_get_params.argtypes = [vp, vp]
    L.tp_mlp_set_params.argtypes = [vp, vp]
    L.tp_mlp_forward.argtypes = [vp, vp, i32, vp]
    L.tp_mlp_loss_grad.argtypes = [vp, vp, vp, i32, C.POINTER(dbl), vp]
    L.tp_mlp_train.argtypes = [vp, vp, vp, i32, i32, i32, C.c_float, C.c_uint64, vp]
    L.tp_mlp_train_log.argtypes = [vp, vp, i32, i32, C.c_float, C.c_uint64, vp]
    L.tp_batch_log.argtypes = [vp, i32]
    L.tp_batch_log_count.argtypes = [vp, vp]
    L.tp_batch_log_read.argtypes = [vp, i32, i32, vp, vp]
    L.tp_batch_set_proposer.argtypes = [vp, vp, i32, dbl]
    L.tp_batch_set_amax.argtypes = [vp, dbl]
    L.tp_batch_set_contraction.argtypes = [vp, i32]
    L.tp_batch_set_family.argtypes = [vp, i32]
    L.tp_anneal_begin.argtypes = [vp, vp, vp, vp, vp, vp]
    L.tp_anneal_steps.argtypes = [vp, vp, vp, vp, vp, vp, i32, dbl, vp, vp, vp, vp]
    L.tp_checkpoint_save.argtypes = [C.c_char_p, vp, vp]
    L.tp_checkpoint_load.argtypes = [C.c_char_p, vp, vp, i64]
    L.tp_load_histograms.argtypes = [vp, i32, i32, i32, i32, vp]
    L.tp_energy_noalloc.argtypes = [vp, vp, vp, vp, i64, i32, C.POINTER(dbl)]
    L.tp_batch_eval_rotations.argtypes = [vp, i32, vp, vp, vp, vp]
    L.tp_batch_set_state.argtypes = [vp, i32, vp, i64, dbl]
    _lib = L
    return L


def check(rc, what=""):
    if rc != 0:
        msg = lib().tp_last_error().decode(errors="replace")
        raise TPError(f"{what} failed (rc={rc}): {msg}")


class PictureParams(C.Structure):
    _fields_ = [
        ("qcut", C.c_double),
        ("envelope_sigma", C.c_double),
        ("lam", C.c_double),
        ("beamstop_width", C.c_double),
        ("photons_per_image", C.c_int),
        ("capacity", C.c_int),
        ("seed", C.c_uint64),
        ("particles", C.c_int),
        ("noise_photons", C.c_int),
        ("noise_sigma", C.c_double),
        ("incident_photon_variance", C.c_int),
        ("beamstop_only", C.c_int),
    ]


class AnnealParams(C.Structure):
    _fields_ = [
        ("nsteps", C.c_int),
        ("ntemp", C.c_int),
        ("initial_stepsize", C.c_double),
        ("stepsizefactor", C.c_double),
        ("initial_temperature_factor", C.c_double),
        ("temperature_decay", C.c_double),
        ("seed", C.c_uint64),
    ]


MAX_LBLOCKS = 32


class AnnealStatus(C.Structure):
    """tp_anneal_status: rotate_all_at_once's state for one L."""
    _fields_ = [
        ("step", C.c_int64),
        ("E", C.c_double),
        ("T", C.c_double),
        ("E0", C.c_double),
        ("T0", C.c_double),
        ("stepsizes", C.c_double * MAX_LBLOCKS),
        ("nl", C.c_int32),
        ("naccepted", C.c_int32),
        ("stopped", C.c_int32),
        ("reserved", C.c_int32),
    ]


class CheckpointInfo(C.Structure):
    """tp_checkpoint_info: header of an annealing checkpoint file."""
    _fields_ = [
        ("params", AnnealParams),
        ("status", AnnealStatus),
        ("min_stepsize", C.c_double),
        ("lam", C.c_double),
        ("dq", C.c_double),
        ("L", C.c_int32),
        ("LMAX", C.c_int32),
        ("N", C.c_int32),
        ("K", C.c_int32),
        ("k3lo", C.c_int32),
        ("k3hi", C.c_int32),
        ("nshell", C.c_int32),
        ("ld", C.c_int32),
    ]


# Entry points declared in include/tp.h (checked by the CPU test suite).
EXPORTS = [
    "tp_abi_version", "tp_last_error", "tp_init", "tp_finalize", "tp_basis_create", "tp_basis_destroy",
    "tp_basis_info", "tp_c3", "tp_ref_create", "tp_ref_destroy", "tp_energy", "tp_state_create",
    "tp_state_destroy", "tp_state_read", "tp_state_energy", "tp_state_c3", "tp_anneal",
    "tp_intensity_from_atoms", "tp_set_timing", "tp_last_timings", "tp_basis_set_precision",
    "tp_basis_get_precision", "tp_energy_batch", "tp_batch_create", "tp_batch_destroy", "tp_batch_autotune",
    "tp_batch_probe", "tp_batch_run", "tp_batch_group", "tp_batch_coeff", "tp_batch_history",
    "tp_batch_eval_quats", "tp_rotate", "tp_mlp_nparam", "tp_mlp_create", "tp_mlp_destroy", "tp_mlp_get_params",
    "tp_mlp_set_params", "tp_mlp_forward", "tp_mlp_loss_grad", "tp_mlp_train", "tp_mlp_train_log", "tp_batch_log",
    "tp_batch_log_count", "tp_batch_log_read", "tp_batch_set_proposer", "tp_batch_set_amax", "tp_batch_set_contraction",
    "tp_batch_set_family", "tp_basis_from_host", "tp_energy_grad",
    "tp_hist_create", "tp_hist_add", "tp_hist_read", "tp_hist_destroy", "tp_two_photons", "tp_retrieve_solution",
    "tp_fit_create", "tp_fit_set_volume", "tp_fit_similarity", "tp_fit_destroy",
    "tp_sh_surface", "tp_sh_coeff", "tp_sph_bessel", "tp_sraar", "tp_pictures", "tp_hist_add_generated",
    "tp_density_surface", "tp_absolute_square", "tp_structure_from_atoms", "tp_two_photons_ewald",
    "tp_retrieve_solution_ewald", "tp_anneal_begin", "tp_anneal_steps", "tp_checkpoint_save", "tp_checkpoint_load",
    "tp_load_histograms", "tp_energy_noalloc", "tp_batch_eval_rotations", "tp_batch_set_state",
]

PREC_FP64, PREC_FP16X3, PREC_BF16, PREC_TRIPLE16, PREC_FP16X3_T64 = 0, 1, 2, 3, 4
MOVE_QUAT, MOVE_PLANE = 0, 1
