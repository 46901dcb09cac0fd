"""ctypes binding of ``libttgpu.so`` (the C ABI in ``include/ttgpu.h``).

The library is loaded from this package directory.  There is no CPU
fallback: if the library is missing or cannot be loaded, every entry point
raises ``NativeLibraryError``.
"""
import ctypes as C
import os
import threading

import numpy as np

from .errors import MissingDataError, TreeTimeError

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libttgpu.so")

TT_OK = 0
TT_EINVAL = -1
TT_ENOMEM_BUDGET = -2
TT_EDEVICE = -3
TT_ERANGE = -4
TT_ESTATE = -5

TT_RECONSTRUCT_TIPS = 1
TT_KEEP_NORMALISERS = 2
TT_PROFILE_MARGINAL = 0
TT_PROFILE_SUBTREE = 1
TT_PROFILE_OUTGROUP = 2
TT_PROFILE_LOG_MSG = 3
TT_BATCH_MARGINAL = 0x00
TT_BATCH_JOINT = 0x10
TT_BATCH_LOGLIK = 0x20
TT_FILL_OVERHANGS = 1
TT_COMPRESS_DEVICE = 2


class NativeLibraryError(RuntimeError):
    """libttgpu.so is absent or failed to load."""


class DeviceError(TreeTimeError):
    """HIP runtime error reported by the engine."""


class BudgetError(TreeTimeError):
    """Admission would exceed the engine's HBM budget."""


class tt_tree_desc(C.Structure):
    _fields_ = [
        ("n_nodes", C.c_int32),
        ("n_states", C.c_int32),
        ("n_patterns", C.c_int32),
        ("n_codes", C.c_int32),
        ("parent", C.POINTER(C.c_int32)),
        ("branch_len", C.POINTER(C.c_double)),
        ("tip_codes", C.POINTER(C.c_uint8)),
        ("code_lut", C.POINTER(C.c_float)),
        ("multiplicity", C.POINTER(C.c_double)),
        ("pi", C.POINTER(C.c_double)),
    ]


_P = C.c_void_p
_I32 = C.c_int32
_U32 = C.c_uint32
_U64 = C.c_uint64
_D = C.c_double
_PD = C.POINTER(C.c_double)
_PF = C.POINTER(C.c_float)
_PI32 = C.POINTER(C.c_int32)
_PU8 = C.POINTER(C.c_uint8)
_PI16 = C.POINTER(C.c_int16)
_I64 = C.c_int64
_PI64 = C.POINTER(C.c_int64)

# name -> (restype, argtypes); every symbol the header declares
SIGNATURES = {
    "tt_last_error": (C.c_char_p, []),
    "tt_version": (_I32, []),
    "tt_engine_create": (_I32, [_I32, _U64, C.POINTER(_P)]),
    "tt_engine_destroy": (_I32, [_P]),
    "tt_engine_usage": (_I32, [_P, C.POINTER(_U64), C.POINTER(_U64)]),
    "tt_tree_create": (_I32, [_P, C.POINTER(tt_tree_desc), C.POINTER(_P)]),
    "tt_tree_destroy": (_I32, [_P]),
    "tt_tree_info": (_I32, [_P, _PI32, _PI32, C.POINTER(_U64)]),
    "tt_tree_pass_shape": (_I32, [_P, _PI32, _PI32, _PI32, _PI32]),
    "tt_tree_set_model": (_I32, [_P, _PD, _PD, _PD, _D]),
    "tt_tree_set_branch_lengths": (_I32, [_P, _PI32, _PD, _I32]),
    "tt_marginal": (_I32, [_P, _U32, _PD, _P]),
    "tt_loglik": (_I32, [_P, _PD, _P]),
    "tt_joint": (_I32, [_P, _U32, _PD, _P]),
    "tt_update_branch": (_I32, [_P, _I32, _D, _PD, _P]),
    "tt_optimize_branch_lengths": (_I32, [_P, _PI16, _I32, _D, _D, _PD, _PI32, _PD, _P]),
    "tt_optimize_branch_lengths_marginal": (_I32, [_P, _I32, _D, _D, _PD, _PI32, _PD, _P]),
    "tt_branch_profile_lh": (_I32, [_P, _I32, _PD, _I32, _PD, _P]),
    "tt_branch_state_counts": (_I32, [_P, _PI16, _I32, _PD, _P]),
    "tt_mutation_counts": (_I32, [_P, _PD, _PD, _PD, _P]),
    "tt_branch_mutation_matrix": (_I32, [_P, _I32, _PD]),
    "tt_get_joint_root": (_I32, [_P, _PD]),
    "tt_joint_from_root": (_I32, [_P, _U32, _PU8, _PD, _P]),
    "tt_get_joint_debug": (_I32, [_P, _PD, _PU8]),
    "tt_get_subtree_prefactor": (_I32, [_P, _PD]),
    "tt_tree_trim": (_I32, [_P]),
    "tt_tree_query": (_I32, [_P]),
    "tt_get_total": (_I32, [_P, _PD]),
    "tt_get_sequence_lh": (_I32, [_P, _PD]),
    "tt_get_states": (_I32, [_P, _PU8]),
    "tt_get_profiles": (_I32, [_P, _I32, _PF]),
    "tt_get_device_buffers": (_I32, [_P, C.POINTER(_PF), C.POINTER(_PU8), C.POINTER(_PD)]),
    "tt_propagate_profile": (_I32, [_P, _I32, _I32, _PD, _I32, _I32, _P, _P]),
    "tt_prob_t_profiles": (_I32, [_P, _P, _P, _I32, _I32, _PD, _I32, _I32, _PD, _P]),
    "tt_prob_t_profiles_noalloc": (_I32, [_P, _P, _P, _I32, _I32, _PD, _I32, _I32, _PD, _P]),
    "tt_batch_marginal": (_I32, [_P, C.POINTER(_P), _I32, _U32, _PD, _P]),
    "tt_batch_info": (_I32, [_P, C.POINTER(_P), _I32, _U32, _PI32, _PI32, _PI32, _PI32, _PD]),
    "tt_tree_set_site_model": (_I32, [_P, _PD, _PD, _PD, _PD, _PD, _PD, _I32, _D]),
    "tt_propagate_profile_site": (_I32, [_P, _I32, _I32, _P, _I32, _I32, _P, _P]),
    "tt_batch_run": (_I32, [_P, C.POINTER(_P), _I32, _U32, _PD, _P]),
    "tt_batch_run_members": (_I32, [_P, C.POINTER(_P), _I32, _U32, _PI32, _I32, _PD, _P]),
    "tt_site_counts": (_I32, [_P, _I32, _PI16, _PD, _PD, _P]),
    "tt_compress_alignment": (_I32, [_P, _P, _I32, _I64, _I64, _I32, _U32, _I64, _PI64, _P, _P, _P, _PI64, _P]),
    "tt_hbm_probe": (_I32, [_I32, _P, _P, _U64, _I32, _PD, _P]),
}

_lib = None
_lock = threading.Lock()


def load():
    """Load the library once; raise NativeLibraryError when it is unavailable."""
    global _lib
    if _lib is not None:
        return _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(LIB_PATH):
            raise NativeLibraryError(
                f"{LIB_PATH} not found: build it with `python -m treetime_amd.build` (there is no CPU fallback)")
        try:
            lib = C.CDLL(LIB_PATH)
        except OSError as e:
            raise NativeLibraryError(f"cannot load {LIB_PATH}: {e}") from e
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc):
    """Map a status code to the reference's exception types."""
    if rc == TT_OK:
        return
    msg = load().tt_last_error().decode(errors="replace")
    if rc == TT_ERANGE:
        raise ValueError(msg)
    if rc == TT_EINVAL:
        raise ValueError(msg)
    if rc == TT_ENOMEM_BUDGET:
        raise BudgetError(msg)
    if rc == TT_ESTATE:
        raise MissingDataError(msg)
    raise DeviceError(msg)


def ptr(a, ctype):
    """ctypes pointer into a C-contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.POINTER(ctype))


def as_c(a, dtype):
    return np.ascontiguousarray(a, dtype=dtype)
