"""ctypes binding of the C ABI in ``include/mqamd.h`` (library ``libmqamd.so``, built in-tree).

This is the only place that loads the HIP library.  There is no fallback: when the library is
missing the import of the simulator backend fails with a message saying how to build it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# MQAMD_LIBRARY names another build of the same library (the host-sanitizer build of _build.py).
LIB_PATH = os.environ.get("MQAMD_LIBRARY") or os.path.join(_HERE, "libmqamd.so")

MQA_OK = 0
MQA_ERR_ARG = 1
MQA_ERR_OOM = 2
MQA_ERR_HIP = 3
MQA_ERR_UNSUPPORTED = 4
MQA_ERR_RUNTIME = 5

MQA_C64 = 0
MQA_C128 = 1

# gate kinds (include/mqamd.h)
GATE_KIND = {
    "I": 0, "X": 1, "Y": 2, "Z": 3, "H": 4, "S": 5, "Sdag": 6, "T": 7, "Tdag": 8, "SX": 9, "SXdag": 10,
    "SWAP": 11, "ISWAP": 12, "ISWAPdag": 13, "MATRIX": 14,
    "RX": 32, "RY": 33, "RZ": 34, "PS": 35, "GP": 36, "Rxx": 37, "Ryy": 38, "Rzz": 39,
    "PMATRIX": 40, "PMATRIX_TAP": 41,
}

# numpy views of the C structs (64 and 32 bytes, natural alignment)
GATE_DTYPE = np.dtype([
    ("kind", "<i4"), ("n_obj", "<i4"), ("obj", "<i4", (6,)), ("ctrl_mask", "<u8"), ("const_val", "<f8"),
    ("term_begin", "<i4"), ("n_term", "<i4"), ("mat_begin", "<i4"), ("requires_grad", "<i4"),
])
TERM_DTYPE = np.dtype([("x_mask", "<u8"), ("y_mask", "<u8"), ("z_mask", "<u8"), ("coeff", "<f8")])
assert GATE_DTYPE.itemsize == 64 and TERM_DTYPE.itemsize == 32


class NativeError(RuntimeError):
    """HIP runtime failure reported by the native library."""


_EXC = {
    MQA_ERR_ARG: ValueError,
    MQA_ERR_OOM: MemoryError,
    MQA_ERR_HIP: NativeError,
    MQA_ERR_UNSUPPORTED: ValueError,
    MQA_ERR_RUNTIME: RuntimeError,
}

_lib = None

_P = ctypes.c_void_p
_PP = ctypes.POINTER(ctypes.c_void_p)
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_U64 = ctypes.c_uint64
_DP = ctypes.POINTER(ctypes.c_double)

_SIGS = {
    "mqa_last_error": ([], ctypes.c_char_p),
    "mqa_version": ([], ctypes.c_char_p),
    "mqa_device_count": ([], ctypes.c_int),
    "mqa_set_option": ([ctypes.c_char_p, _I64], ctypes.c_int),
    "mqa_get_option": ([ctypes.c_char_p, ctypes.POINTER(_I64)], ctypes.c_int),
    "mqa_circuit_create": ([_P, _I32, _P, _P, _I32, _P, _I64, _I32, _P, _PP], ctypes.c_int),
    "mqa_circuit_destroy": ([_P], ctypes.c_int),
    "mqa_circuit_set_custom_matrices": ([_P, _I64, _I64, _P], ctypes.c_int),
    "mqa_circuit_set_tile_qubits": ([_P, _I32, _I32, _I32], ctypes.c_int),
    "mqa_circuit_get_tile_qubits": ([_P, _I32, _I32, ctypes.POINTER(_I32)], ctypes.c_int),
    "mqa_circuit_schedule": ([_P, _I32, _I32, _I32, _P, _I64, ctypes.POINTER(_I64)], ctypes.c_int),
    "mqa_circuit_jit_compile": ([_P, _I32, _I32, _I32, ctypes.POINTER(ctypes.c_double)], ctypes.c_int),
    "mqa_ham_create": ([_P, _I32, _PP], ctypes.c_int),
    "mqa_ham_create_csr": ([_U64, _U64, _P, _P, _P, _PP], ctypes.c_int),
    "mqa_ham_destroy": ([_P], ctypes.c_int),
    "mqa_sim_create": ([_I32, _I32, ctypes.c_uint32, _PP], ctypes.c_int),
    "mqa_sim_destroy": ([_P], ctypes.c_int),
    "mqa_sim_reset": ([_P], ctypes.c_int),
    "mqa_sim_copy": ([_P, _PP], ctypes.c_int),
    "mqa_sim_n_qubits": ([_P, ctypes.POINTER(_I32), ctypes.POINTER(_I32)], ctypes.c_int),
    "mqa_sim_get_qs": ([_P, _P, _U64], ctypes.c_int),
    "mqa_sim_set_qs": ([_P, _P, _U64], ctypes.c_int),
    "mqa_sim_apply_circuit": ([_P, _P, _P, _I32], ctypes.c_int),
    "mqa_sim_apply_dense_blocks": ([_P, _I32, _P, _P, _P], ctypes.c_int),
    "mqa_sim_apply_hamiltonian": ([_P, _P], ctypes.c_int),
    "mqa_sim_expectation": ([_P, _P, _P, _P, _I32, _P], ctypes.c_int),
    "mqa_sim_expectation_lr": ([_P, _P, _P, _P, _P, _P, _I32, _P], ctypes.c_int),
    "mqa_sim_expectation_noalloc": ([_P, _P, _P], ctypes.c_int),
    "mqa_sim_grad_batch": ([_P, _P, _I32, _P, _P, _I64, _I32, _P, _I32, _P, _P], ctypes.c_int),
    "mqa_sim_grad_nonherm": ([_P, _P, _P, _I32, _P, _P, _P, _I64, _I32, _P, _I32, _P, _P], ctypes.c_int),
    "mqa_sim_sample": ([_P, _P, _P, _I32, _P, _U64, _P], ctypes.c_int),
    "mqa_sim_expectation_angles": ([_P, _P, _I32, _P, _P, _I64, _P], ctypes.c_int),
    "mqa_sim_astype": ([_P, _I32, _PP], ctypes.c_int),
    "mqa_grad_items": ([_I32, _P, _P, _P, _P, _P, _I32, _P, _P], ctypes.c_int),
    "mqa_grad_workspace_plan": ([_U64, _U64, _U64, _I32, _I32, _I64, _P, ctypes.POINTER(_I64), ctypes.POINTER(_I32),
                                 _P, _P], ctypes.c_int),
    "mqa_inner_product": ([_P, _P, _P], ctypes.c_int),
    "mqa_profile_get": ([ctypes.c_char_p, ctypes.POINTER(_I64), _DP, _DP], ctypes.c_int),
    "mqa_profile_reset": ([], ctypes.c_int),
    "mqa_copy_probe": ([_U64, _I32, _I32, _DP], ctypes.c_int),
}


def exported_symbols():
    """Names of every entry point the binding declares (== the functions of include/mqamd.h)."""
    return sorted(_SIGS)


def lib():
    """Load ``libmqamd.so`` once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build the HIP library first (python -m mindquantum_amd._build)"
        )
    handle = ctypes.CDLL(LIB_PATH)
    for name, (args, res) in _SIGS.items():
        fn = getattr(handle, name)
        fn.argtypes = args
        fn.restype = res
    _lib = handle
    return _lib


def check(status):
    """Raise the Python exception matching a non-zero status."""
    if status != MQA_OK:
        msg = lib().mqa_last_error().decode("utf-8", "replace")
        raise _EXC.get(status, RuntimeError)(msg)


def ptr(arr):
    """Data pointer of a contiguous numpy array (None for None)."""
    if arr is None:
        return None
    return ctypes.c_void_p(arr.ctypes.data)


def grad_workspace_plan(budget, state_bytes, side_bytes, n_hams, n_checkpoints, batch, held):
    """(chunk, keep, need[4], release[4]) of mqa_grad_workspace_plan (host only)."""
    h = np.ascontiguousarray(held, dtype=np.uint64)
    need = np.zeros(4, dtype=np.uint64)
    rel = np.zeros(4, dtype=np.int32)
    chunk, keep = _I64(0), _I32(0)
    check(lib().mqa_grad_workspace_plan(int(budget), int(state_bytes), int(side_bytes), int(n_hams),
                                        int(n_checkpoints), int(batch), ptr(h), ctypes.byref(chunk),
                                        ctypes.byref(keep), ptr(need), ptr(rel)))
    return int(chunk.value), bool(keep.value), [int(x) for x in need], [bool(x) for x in rel]


def profile(kernel):
    """(launches, total milliseconds, algorithmic bytes) of a kernel family since the last reset."""
    n, ms, b = _I64(0), ctypes.c_double(0), ctypes.c_double(0)
    check(lib().mqa_profile_get(kernel.encode(), ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)))
    return int(n.value), float(ms.value), float(b.value)


def profile_reset():
    check(lib().mqa_profile_reset())


def device_count():
    return int(lib().mqa_device_count())


def set_option(key, value):
    check(lib().mqa_set_option(key.encode(), int(value)))


def get_option(key):
    v = _I64(0)
    check(lib().mqa_get_option(key.encode(), ctypes.byref(v)))
    return int(v.value)
