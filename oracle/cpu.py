"""TEST INFRASTRUCTURE ONLY: ctypes binding of the C++ oracle (oracle/mqvector_oracle.cpp).

The oracle restates the reference CPU mqvector algorithm; see the header of the .cpp file for
the reference file:line each routine follows.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# ORACLE_LIBRARY names another build of the same source (`make sanitize`: ASan/UBSan).
LIB = os.environ.get("ORACLE_LIBRARY") or os.path.join(HERE, "_build", "liboracle.so")
_lib = None


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = ctypes.CDLL(LIB)
        P = ctypes.c_void_p
        I = ctypes.c_int
        _lib.oracle_last_error.restype = ctypes.c_char_p
        _lib.oracle_apply_circuit.argtypes = [I, I, P, P, I, P, P, P, P]
        _lib.oracle_apply_hamiltonian.argtypes = [I, I, P, P, P, I]
        _lib.oracle_expectation.argtypes = [I, I, P, P, P, I, P]
        _lib.oracle_vdot.argtypes = [I, I, P, P, P]
        _lib.oracle_grad_batch.argtypes = [I, I, P, P, I, P, P, P, P, P, I, P, ctypes.c_int64, I, P, I, I, P]
        _lib.oracle_num_threads.restype = I
        _lib.oracle_set_num_threads.argtypes = [I]
        _lib.oracle_set_worker_threads.argtypes = [I]
    return _lib


def _p(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _chk(st):
    if st != 0:
        raise RuntimeError(lib().oracle_last_error().decode())


def _code(dtype):
    return 0 if np.dtype(dtype) == np.complex64 else 1


def zero_state(n, dtype=np.complex128):
    qs = np.zeros(1 << n, dtype=dtype)
    qs[0] = 1
    return qs


def apply_circuit(qs, ir_arrays, params):
    """Apply a lowered circuit (mindquantum_amd.ir.lower output) to ``qs`` in place."""
    gates, tp, tc, mats, _, _ = ir_arrays
    n = int(np.log2(qs.size))
    params = np.ascontiguousarray(params, dtype=np.float64)
    _chk(lib().oracle_apply_circuit(n, _code(qs.dtype), _p(qs), _p(gates), len(gates), _p(tp), _p(tc), _p(mats),
                                    _p(params) if params.size else None))
    return qs


def apply_hamiltonian(qs, terms):
    out = np.zeros_like(qs)
    n = int(np.log2(qs.size))
    _chk(lib().oracle_apply_hamiltonian(n, _code(qs.dtype), _p(qs), _p(out), _p(terms), len(terms)))
    return out


def expectation(bra, ket, terms):
    out = np.zeros(2)
    n = int(np.log2(ket.size))
    _chk(lib().oracle_expectation(n, _code(ket.dtype), _p(bra), _p(ket), _p(terms), len(terms), _p(out)))
    return complex(out[0], out[1])


def vdot(a, b):
    out = np.zeros(2)
    n = int(np.log2(a.size))
    _chk(lib().oracle_vdot(n, _code(a.dtype), _p(a), _p(b), _p(out)))
    return complex(out[0], out[1])


def grad_batch(n, dtype, ir_arrays, hams_terms, enc, ans, batch_threads=1, init=None, worker_threads=0):
    """Reference hermitian adjoint gradient: returns f [b, H] and g [b, H, P] (complex).

    batch_threads worker threads share the samples (gradient.h:270-299, one sample runs inline); each
    worker uses worker_threads OpenMP threads (0: the OpenMP default)."""
    gates, tp, tc, mats, _, _ = ir_arrays
    enc = np.ascontiguousarray(enc, dtype=np.float64)
    ans = np.ascontiguousarray(ans, dtype=np.float64)
    if enc.ndim == 1:
        enc = enc.reshape(1, -1)
    batch, n_enc = enc.shape
    n_ans = ans.size
    terms = np.concatenate(hams_terms) if hams_terms else None
    off = np.zeros(len(hams_terms) + 1, dtype=np.int32)
    for i, t in enumerate(hams_terms):
        off[i + 1] = off[i] + len(t)
    n_params = n_enc + n_ans
    out = np.zeros((batch, len(hams_terms), 1 + n_params, 2), dtype=np.float64)
    if init is not None:
        init = np.ascontiguousarray(init, dtype=dtype)
    lib().oracle_set_worker_threads(int(worker_threads))
    _chk(lib().oracle_grad_batch(n, _code(dtype), _p(init), _p(gates), len(gates), _p(tp), _p(tc), _p(mats),
                                 _p(terms), _p(off), len(hams_terms), _p(enc) if n_enc else None, batch, n_enc,
                                 _p(ans) if n_ans else None, n_ans, batch_threads, _p(out)))
    res = out[..., 0] + 1j * out[..., 1]
    return res[:, :, 0], res[:, :, 1:]


def set_num_threads(n):
    lib().oracle_set_num_threads(int(n))


def num_threads():
    return int(lib().oracle_num_threads())
