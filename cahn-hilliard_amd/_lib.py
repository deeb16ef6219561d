"""ctypes bindings of include/ch_hip.h (libch_hip.so, built for gfx950).

There is no CPU fallback: if the shared library is missing, or no GPU is
visible, the calls raise.
"""
import ctypes
import os
import re

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libch_hip.so")
HEADER = os.path.join(os.path.dirname(_HERE), "include", "ch_hip.h")

CH_OK, CH_EINVAL, CH_EHIP, CH_ENAN, CH_EUNSUPPORTED, CH_ENOMEM, CH_ESTATE = range(7)
FP64, FP32 = 0, 1
EULER, EULER_MOBILITY, PSEUDOSPECTRAL, PSEUDOSPECTRAL_MOBILITY = 0, 1, 2, 3
LANDAU, SIMPLE_WERTHEIM = 0, 1
MOB_CONSTANT, MOB_FREE_ENERGY, MOB_REGULARISED, MOB_LINEAR = 0, 1, 2, 3
# parameter-gradient slots of ch_sys_adjoint / ch_sys_free_energy_grad
GRAD_EPSILON, GRAD_K, GRAD_M, GRAD_DELTA, GRAD_B2, NGRADS = 0, 1, 2, 3, 4, 5

SALEH_WERTHEIM, RICCI_WERTHEIM, GENERIC_WERTHEIM = 2, 3, 4
MODELS_SPECIES = {"saleh": SALEH_WERTHEIM, "ricci": RICCI_WERTHEIM, "generic_wertheim": GENERIC_WERTHEIM}
INTEGRATORS = {"euler": EULER, "euler_mobility": EULER_MOBILITY, "pseudospectral": PSEUDOSPECTRAL,
               "pseudospectral_mobility": PSEUDOSPECTRAL_MOBILITY}
MODELS = {"landau": LANDAU, "simple_wertheim": SIMPLE_WERTHEIM}
MOBILITIES = {"constant": MOB_CONSTANT, "free_energy": MOB_FREE_ENERGY, "regularised": MOB_REGULARISED,
              "linear": MOB_LINEAR}


class ChDesc(ctypes.Structure):
    _fields_ = [("dims", ctypes.c_int32), ("N", ctypes.c_int32), ("precision", ctypes.c_int32),
                ("integrator", ctypes.c_int32), ("model", ctypes.c_int32), ("mobility", ctypes.c_int32),
                ("valence", ctypes.c_int32), ("use_dealias", ctypes.c_int32),
                ("epsilon", ctypes.c_double), ("B2", ctypes.c_double), ("delta", ctypes.c_double),
                ("regularisation_delta", ctypes.c_double), ("M", ctypes.c_double), ("rho_min", ctypes.c_double),
                ("rho_ref", ctypes.c_double), ("dt", ctypes.c_double), ("dx", ctypes.c_double),
                ("k", ctypes.c_double), ("S", ctypes.c_double), ("with_noise", ctypes.c_int32),
                ("reserved0", ctypes.c_int32), ("noise_rescale_factor", ctypes.c_double), ("seed", ctypes.c_uint64),
                ("use_gmres", ctypes.c_int32), ("gmres_restart", ctypes.c_int32), ("gmres_max_iter", ctypes.c_int32),
                ("reserved1", ctypes.c_int32), ("gmres_tol", ctypes.c_double),
                ("density_conversion", ctypes.c_double)]


MAX_SPECIES, MAX_PATCHES = 8, 16
OBS_QUEUE = 8  # CH_OBS_QUEUE


class ChSpeciesModel(ctypes.Structure):
    """ch_species_model: a multi-species Wertheim model (include/ch_hip.h)."""
    _fields_ = [("model", ctypes.c_int32), ("n_species", ctypes.c_int32), ("k", ctypes.c_double * MAX_SPECIES),
                ("saleh_valence", ctypes.c_int32 * 3), ("reserved0", ctypes.c_int32),
                ("saleh_B2", ctypes.c_double), ("saleh_B3", ctypes.c_double), ("saleh_delta_AA", ctypes.c_double),
                ("saleh_delta_BB", ctypes.c_double), ("ricci_B2", ctypes.c_double), ("ricci_delta_00", ctypes.c_double),
                ("ricci_delta_12", ctypes.c_double), ("n_patches", ctypes.c_int32), ("reserved1", ctypes.c_int32),
                ("patch_count", (ctypes.c_int32 * MAX_PATCHES) * MAX_SPECIES),
                ("delta", (ctypes.c_double * MAX_PATCHES) * MAX_PATCHES),
                ("B2", (ctypes.c_double * MAX_SPECIES) * MAX_SPECIES)]


class CHError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"ch status {status}: {msg}")
        self.status = status


def header_functions(path=HEADER):
    """Names of every function the public header declares."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ch_[a-z0-9_]+)\s*\(", text)))


_lib = None


def lib():
    global _lib
    if _lib is not None:
        return _lib
    # torch ships its own libamdhip64.so.7 (same SONAME as /opt/rocm's).  Loading
    # it first makes this library bind to that runtime, so the process holds one
    # HIP/HSA runtime; loading ours first would make torch open a second one,
    # which then sees no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(there is no CPU fallback)")
    L = ctypes.CDLL(LIB_PATH)
    P, I64, I32, D = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_double
    DP = ctypes.POINTER(ChDesc)
    PP = ctypes.POINTER(ctypes.c_void_p)
    DD = ctypes.POINTER(ctypes.c_double)
    sigs = {
        "ch_abi_version": (ctypes.c_int, []),
        "ch_status_string": (ctypes.c_char_p, [ctypes.c_int]),
        "ch_ctx_create": (ctypes.c_int, [ctypes.c_int, PP]),
        "ch_ctx_destroy": (None, [P]),
        "ch_ctx_last_error": (ctypes.c_char_p, [P]),
        "ch_ctx_synchronize": (ctypes.c_int, [P]),
        "ch_ctx_stream": (P, [P]),
        "ch_sys_create": (ctypes.c_int, [P, DP, PP]),
        "ch_sys_create_species": (ctypes.c_int, [P, DP, P, PP]),
        "ch_sys_species": (ctypes.c_int32, [P]),
        "ch_sys_destroy": (None, [P]),
        "ch_sys_upload": (ctypes.c_int, [P, P]),
        "ch_sys_download": (ctypes.c_int, [P, P]),
        "ch_sys_upload_device": (ctypes.c_int, [P, P]),
        "ch_sys_download_device": (ctypes.c_int, [P, P]),
        "ch_sys_download_mobility": (ctypes.c_int, [P, P]),
        "ch_sys_evolve": (ctypes.c_int, [P, I64]),
        "ch_sys_steps_per_launch": (ctypes.c_int32, [P]),
        "ch_sys_region_shape": (ctypes.c_int32, [P]),
        "ch_sys_gmres_iterations": (ctypes.c_int32, [P]),
        "ch_sys_observables": (ctypes.c_int, [P, DD, DD]),
        "ch_sys_observables_queue": (ctypes.c_int, [P]),
        "ch_sys_observables_fetch": (ctypes.c_int, [P, DD, DD]),
        "ch_sys_observables_pending": (ctypes.c_int32, [P]),
        "ch_sys_evolve_observe": (ctypes.c_int, [P, I64]),
        "ch_observables_device_noalloc": (ctypes.c_int, [P, DP, P, DD, DD]),
        "ch_sq_nshells": (I32, [I32]),
        "ch_sys_structure_factor": (ctypes.c_int, [P, P, I32]),
        "ch_sys_sq_loss": (ctypes.c_int, [P, P, I32, DD, P]),
        "ch_batch_create": (ctypes.c_int, [P, I32, DP, PP]),
        "ch_batch_create_ex": (ctypes.c_int, [P, I32, DP, I32, PP]),
        "ch_batch_compensation": (I32, [P]),
        "ch_batch_destroy": (None, [P]),
        "ch_batch_upload": (ctypes.c_int, [P, I32, P]),
        "ch_batch_download": (ctypes.c_int, [P, I32, P]),
        "ch_batch_evolve": (ctypes.c_int, [P, I64]),
        "ch_batch_observables": (ctypes.c_int, [P, P, P]),
        "ch_batch_steps_per_launch": (ctypes.c_int32, [P]),
        "ch_batch_autotune": (ctypes.c_int, [P, ctypes.POINTER(I32), ctypes.POINTER(I32)]),
        "ch_sys_record": (ctypes.c_int, [P, I64, I32]),
        "ch_sys_adjoint": (ctypes.c_int, [P, P, P, P]),
        "ch_sys_adjoint_device": (ctypes.c_int, [P, P, P, P]),
        "ch_sys_free_energy_grad": (ctypes.c_int, [P, DD, P, P]),
        "ch_der_bulk_free_energy": (ctypes.c_int, [P, DP, P, P, I64]),
        "ch_set_mobility": (ctypes.c_int, [P, DP, P, P, I64]),
    }
    for name, (res, args) in sigs.items():
        if hasattr(L, name):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
    _lib = L
    return L
