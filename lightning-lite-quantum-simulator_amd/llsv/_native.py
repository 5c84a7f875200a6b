"""ctypes binding of libllsv.so (the C ABI of include/llsv.h).

The library is built in-tree by build_native.py.  Loading fails loudly when it is
missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libllsv.so")

LLSV_OK = 0
LLSV_ERR_ARG = -1
LLSV_ERR_NOMEM = -2
LLSV_ERR_DEVICE = -3
LLSV_ERR_HANDLE = -4
LLSV_ERR_NUMERIC = -5

_P = ctypes.c_void_p
_I = ctypes.c_int
_U64 = ctypes.c_uint64
_D = ctypes.c_double

# name -> (restype, argtypes); mirrors include/llsv.h
SIGNATURES = {
    "llsv_create": (_I, [_I, ctypes.POINTER(_P)]),
    "llsv_create_from_host": (_I, [_P, _U64, ctypes.POINTER(_P)]),
    "llsv_destroy": (_I, [_P]),
    "llsv_reset": (_I, [_P]),
    "llsv_num_qubits": (_I, [_P]),
    "llsv_size": (_U64, [_P]),
    "llsv_apply_1q": (_I, [_P, _I, _P]),
    "llsv_apply_cnot": (_I, [_P, _I, _I]),
    "llsv_apply_cz": (_I, [_P, _I, _I]),
    "llsv_apply_mcz": (_I, [_P, _P, _I]),
    "llsv_apply_dense": (_I, [_P, _I, _P, _P]),
    "llsv_apply_program": (_I, [_P, _P, _U64]),
    "llsv_plan_passes": (_I, [_I, _P, _U64, _I, ctypes.POINTER(_U64), _P, _P]),
    "llsv_program_tile_bits": (_I, [_P]),
    "llsv_set_program_jit": (_I, [_I]),
    "llsv_program_jit_stats": (None, [ctypes.POINTER(_U64), ctypes.POINTER(_U64), ctypes.POINTER(_U64)]),
    "llsv_program_jit_precompile": (_I, [_I, _P, _U64]),
    "llsv_program_jit_wait": (None, []),
    "llsv_norm2": (_I, [_P, ctypes.POINTER(_D)]),
    "llsv_expect_z": (_I, [_P, _I, ctypes.POINTER(_D)]),
    "llsv_probabilities": (_I, [_P, _P, _U64]),
    "llsv_expect_pauli": (_I, [_P, _P, _I, _P, _P]),
    "llsv_sample": (_I, [_P, _U64, _U64, _P]),
    "llsv_sample_noalloc": (_I, [_P, _U64, _U64, _P]),
    "llsv_copy_to_host": (_I, [_P, _P, _U64]),
    "llsv_copy_from_host": (_I, [_P, _P, _U64]),
    "llsv_sync": (_I, [_P]),
    "llsv_copy_state": (_I, [_P, _P]),
    "llsv_stream_sync": (_I, [_P]),
    "llsv_pc_workspace_bytes": (_U64, [_I, _I, _I]),
    "llsv_pc_energy": (_I, [_I, _I, _P, _I, _I, _P, _P, _I, _P, _P, _U64, _P]),
    "llsv_noisy_create": (_I, [ctypes.c_double, _U64, _I, ctypes.POINTER(_P)]),
    "llsv_noisy_destroy": (None, [_P]),
    "llsv_noisy_reseed": (_I, [_P, _U64]),
    "llsv_noisy_site": (_I, [_P, _I, ctypes.POINTER(ctypes.c_uint32)]),
    "llsv_noisy_run": (_I, [_P, _I, _P, _I, _I, ctypes.c_int64, _P, _P, _P]),
    "llsv_zgemm": (_I, [_I, _I, _I, _P, _P, _P]),
    "llsv_zgemm_time": (_I, [_I, _I, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "llsv_expm": (_I, [_I, _P, _P]),
    "llsv_lplan_create": (_I, [_I, _I, ctypes.c_int64, ctypes.c_double, _I, _P, _I, _P, ctypes.POINTER(_P)]),
    "llsv_lplan_create_h": (_I, [_I, _I, ctypes.c_int64, ctypes.c_double, _I, _P, _I, _P, ctypes.POINTER(_P)]),
    "llsv_lplan_destroy": (None, [_P]),
    "llsv_lplan_run": (_I, [_P, _P, ctypes.c_int64, ctypes.c_int64, _U64, _I, _P, _P]),
    "llsv_pc_precompile": (_I, [_I, _P, _I, ctypes.POINTER(_I)]),
    "llsv_pc_adjoint": (_I, [_I, _I, _P, _I, _I, _P, _P, _I, _P, _P, _P, _U64, _P]),
    "llsv_traj_run": (_I, [_P, _I, _P, _I, _P, _P, _I, _U64, _P, _P, _P, _P]),
    "llsv_traj_plan_create": (_I, [_P, _I, _P, _I, _P, _P, _I, _I, ctypes.POINTER(_P)]),
    "llsv_traj_plan_run": (_I, [_P, _U64, _P, _P, _P]),
    "llsv_traj_plan_config": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_D)]),
    "llsv_traj_plan_destroy": (_I, [_P]),
    "llsv_traj_default_config": (_I, [_I, _I, _I]),
    "llsv_traj_tune_load": (_I, [ctypes.c_char_p]),
    "llsv_traj_tune_save": (_I, [ctypes.c_char_p]),
    "llsv_traj_tune_entries": (_I, []),
    "llsv_mark": (_I, [_P, _I]),
    "llsv_mark_elapsed": (_I, [_P, _I, _I, ctypes.POINTER(ctypes.c_float)]),
    "llsv_device_ptr": (_P, [_P]),
    "llsv_stream": (_P, [_P]),
    "llsv_timer_start": (_I, [_P]),
    "llsv_timer_stop": (_I, [_P, ctypes.POINTER(ctypes.c_float)]),
    "llsv_test_fail_allocation": (ctypes.c_long, [ctypes.c_long]),
    "llsv_last_error": (ctypes.c_char_p, []),
    "llsv_version": (ctypes.c_char_p, []),
}


class LlsvOp(ctypes.Structure):
    """llsv_op of include/llsv.h."""

    _fields_ = [
        ("kind", ctypes.c_int32),
        ("nq", ctypes.c_int32),
        ("q", ctypes.c_int32 * 8),
        ("m", ctypes.c_double * 8),
        ("dense", ctypes.c_void_p),
    ]


_lib = None


def lib():
    """The loaded libllsv.so with argtypes set.  Raises if it was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"libllsv.so not built ({LIB_PATH}); run build_native.py")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def hip_runtimes_loaded():
    """Paths of the distinct HIP runtime libraries mapped into this process.  torch ships its
    own copy; when torch is imported before libllsv both resolve to one runtime (same soname)
    and torch's streams can be handed to libllsv.  With two runtimes they cannot."""
    paths = set()
    try:
        with open("/proc/self/maps") as f:
            for line in f:
                if "libamdhip64" in line:
                    paths.add(os.path.realpath(line.split()[-1]))
    except OSError:
        pass
    return sorted(paths)


class LlsvError(RuntimeError):
    pass


def check(rc: int, what: str = ""):
    if rc == LLSV_OK:
        return
    msg = lib().llsv_last_error().decode(errors="replace")
    if rc == LLSV_ERR_ARG:
        raise ValueError(msg or what)
    if rc == LLSV_ERR_NOMEM:
        raise MemoryError(msg or what)
    raise LlsvError(msg or what)
