"""ctypes binding of the C ABI declared in include/gpt_amd_wc.h.

The product has no CPU fallback: if the HIP library is missing, importing the
operator layer raises.  Build it with ``python -m gpt_amd.build``.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
# GPT_AMD_LIB selects another build of the same library (e.g. a tuning variant)
LIB_PATH = os.environ.get("GPT_AMD_LIB", os.path.join(_HERE, "libgpt_amd_wc.so"))

WC_OK, WC_ERR_ARG, WC_ERR_HIP, WC_ERR_NUMERIC = 0, 1, 2, 3
WC_DOUBLE, WC_SINGLE = 0, 1
WC_STORAGE_NATIVE, WC_STORAGE_HALF, WC_STORAGE_BF16 = 0, 2, 3
# inner_storage names of the Python API -> the C codes
STORAGE_CODES = {None: WC_STORAGE_NATIVE, "half": WC_STORAGE_HALF, "bf16": WC_STORAGE_BF16}


class wc_params(ctypes.Structure):
    _fields_ = [
        ("fdims", ctypes.c_int * 4),
        ("mass", ctypes.c_double),
        ("csw_r", ctypes.c_double),
        ("csw_t", ctypes.c_double),
        ("cF", ctypes.c_double),
        ("xi_0", ctypes.c_double),
        ("nu", ctypes.c_double),
        ("isAnisotropic", ctypes.c_int),
        ("boundary_phases", (ctypes.c_double * 2) * 4),
        ("precision", ctypes.c_int),
        ("compress_links", ctypes.c_int),
    ]


class wc_solver_stats(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_long),
        ("converged", ctypes.c_int),
        ("r2", ctypes.c_double),
        ("b2", ctypes.c_double),
    ]


# every symbol the header declares, with its ctypes signature
_P = ctypes.c_void_p
_I, _L, _D = ctypes.c_int, ctypes.c_long, ctypes.c_double
SIGNATURES = {
    "wc_create": (_I, [ctypes.POINTER(wc_params), _P, _P, ctypes.POINTER(_P)]),
    "wc_update": (_I, [_P, _P, _P]),
    "wc_set_mass": (_I, [_P, _D, _P]),
    "wc_destroy": (_I, [_P]),
    "wc_half_volume": (_L, [_P]),
    "wc_precision": (_I, [_P]),
    "wc_apply": (_I, [_P, _I, _P, _P, _I, _P]),
    "wc_axpy": (_I, [_I, _L, _D, _D, _P, _P, _P, _P]),
    "wc_scale": (_I, [_I, _L, _D, _D, _P, _P, _P]),
    "wc_norm2": (_I, [_I, _L, _P, ctypes.POINTER(_D), _P]),
    "wc_inner_product": (_I, [_I, _L, _P, _P, ctypes.POINTER(_D), _P]),
    "wc_axpy_norm2": (_I, [_I, _L, _D, _D, _P, _P, _P, ctypes.POINTER(_D), _P]),
    "wc_convert": (_I, [_I, _P, _I, _P, _L, _P]),
    "wc_field_import": (_I, [ctypes.POINTER(_I), _I, _P, _P, _P]),
    "wc_field_export": (_I, [ctypes.POINTER(_I), _I, _P, _P, _P]),
    "wc_cg_eo2_ne": (_I, [_P, _I, _P, _P, _D, _L, _P, ctypes.POINTER(wc_solver_stats), _P]),
    "wc_solve_eo2_ne": (_I, [_P, _I, _P, _P, _D, _L, _P, ctypes.POINTER(wc_solver_stats), _P]),
    "wc_apply_mpc": (_I, [_P, _I, _I, _P, _P, _P]),
    "wc_apply_block": (_I, [_P, _I, _I, _P, _P, _I, _P]),
    "wc_apply_mpc_block": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "wc_cg_eo2_ne_block": (_I, [_P, _I, _I, _P, _P, _D, _L, _P, _P, _P, _P, _P, _P]),
    "wc_solve_eo2_ne_block": (_I, [_P, _I, _I, _P, _P, _D, _L, _P, _P, _P, _P, _P, _P]),
    "wc_solve_eo2_ne_mixed": (_I, [_P, _P, _I, _I, _P, _P, _D, _L, _D, _I, _P, _P, _P, _P, _P]),
    "wc_tune_hop": (_I, [_P, _I, _I, _P, _P, _P]),
    "wc_set_hop_shape": (_I, [_P, _I, _P]),
    "wc_get_hop_shape": (_I, [_P, _I, _P]),
    "wc_set_iteration_chunk": (_I, [_P, _I]),
    "wc_trim": (_I, [_P]),
    "wc_release_cached_memory": (_I, [_P]),
    "wc_gauge_observables": (_I, [_P, _P, _P, _P, _P]),
    "wc_halo_pack": (_I, [_P, _I, _I, _I, _P, _P, _P, _P]),
    "wc_halo_unpack": (_I, [_P, _I, _I, _I, _P, _P, _P, _P]),
    "wc_bandwidth_probe": (_I, [_I, _P, _P, _L, _P]),
    "wc_halo_interior": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "wc_block_cg_eo2_ne": (_I, [_P, _I, _I, _P, _P, _D, _D, _L, _L, _P, _P, _P, _P]),
    "wc_solve_eo2_ne_block_cg": (_I, [_P, _I, _I, _P, _P, _D, _D, _L, _L, _P, _P, _P, _P]),
    "wc_iteration_runs": (_I, [_P, _P, _P]),
    "wc_batch_create": (_I, [_P, _P, _I, _I, _D, _I, _P, ctypes.POINTER(_P)]),
    "wc_batch_destroy": (_I, [_P]),
    "wc_batch_admit": (_I, [_P, _I, _P, _P, _D, _L]),
    "wc_batch_admit_mass": (_I, [_P, _I, _D, _P, _P, _D, _L]),
    "wc_batch_step": (_I, [_P, _I, _P, _P]),
    "wc_batch_extract": (_I, [_P, _I, _P, _P, _P, _P]),
    "wc_batch_slots": (_I, [_P, _P, _P]),
    "wc_batch_set_mass_sets": (_I, [_P, _I]),
    "wc_set_hop_mma": (_I, [_P, _I]),
    "wc_get_hop_mma": (_I, [_P, _I, _P]),
    "wc_apply_block_s16": (_I, [_P, _I, _I, _I, _P, _P, _I, _P]),
    "wc_apply_mpc_block_s16": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "wc_apply_mpc_block_noalloc": (_I, [_P, _I, _I, _I, _P, _P, _P]),
    "wc_last_error": (ctypes.c_char_p, []),
}

_lib = None


def load():
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"gpt_amd: HIP library {LIB_PATH} is missing; build it with `python -m gpt_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    if rc != WC_OK:
        msg = load().wc_last_error()
        raise RuntimeError(msg.decode() if msg else f"gpt_amd error {rc}")
