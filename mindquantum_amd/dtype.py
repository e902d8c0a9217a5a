"""Precision names (mindquantum/dtype/dtype.py:44-263, reduced to the two state types)."""
import numpy as np

complex64 = np.complex64  # pylint: disable=invalid-name
complex128 = np.complex128  # pylint: disable=invalid-name
float32 = np.float32  # pylint: disable=invalid-name
float64 = np.float64  # pylint: disable=invalid-name

_ALIASES = {
    "complex64": np.complex64, "c64": np.complex64, "float32": np.complex64, "float": np.complex64,
    "complex128": np.complex128, "c128": np.complex128, "float64": np.complex128, "double": np.complex128,
}


def to_mq_type(dtype):
    """Normalise a dtype spec to numpy complex64 / complex128."""
    if dtype is None:
        return np.complex128
    if isinstance(dtype, str):
        if dtype.lower() not in _ALIASES:
            raise TypeError(f"Unsupported dtype {dtype}")
        return _ALIASES[dtype.lower()]
    d = np.dtype(dtype)
    if d in (np.complex64, np.float32):
        return np.complex64
    if d in (np.complex128, np.float64):
        return np.complex128
    raise TypeError(f"Unsupported dtype {dtype}")


def precision_str(dtype):
    return "single precision" if to_mq_type(dtype) == np.complex64 else "double precision"


def is_same_precision(a, b):
    return to_mq_type(a) == to_mq_type(b)


def native_code(dtype):
    """C ABI precision code (MQA_C64 / MQA_C128)."""
    return 0 if to_mq_type(dtype) == np.complex64 else 1


def real_type(dtype):
    return np.float32 if to_mq_type(dtype) == np.complex64 else np.float64
