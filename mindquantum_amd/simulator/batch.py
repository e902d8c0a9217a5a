"""Expectation and gradient of many independent circuits in shared launches.

The reference evaluates a list of circuits (for instance one QAOA instance per graph) with one
``Simulator`` and one ``get_expectation_with_grad`` call per circuit (mqsim.py:255-439).  Here
the whole list goes to ``mqa_grad_items``: every circuit starts from |0...0> on its own qubit
count, and each sweep step of all circuits runs as one launch over all of their tiles, so small
circuits do not leave the GPU idle.
"""
import ctypes

import numpy as np

from .. import _native
from .. import dtype as mqdtype
from ..core.operators import Hamiltonian
from ..ir import NativeCache

# Native circuits (and with them their schedules, device plans and compiled passes) of the circuits
# a caller passes again: held weakly per circuit object and revalidated after any mutation
# (ir.NativeCache); the bound keeps a caller that builds fresh circuits per call from holding
# more than this many.
_NATIVE_CACHE = NativeCache(1024)


def _native_circuit(circ, names):
    return _NATIVE_CACHE.get(circ, names)


def expectation_with_grad_batch(circuits, hamiltonians, params, n_qubits=None, dtype=None):
    """Return ``(f, grads)`` for ``len(circuits)`` independent items.

    Args:
        circuits: list of Circuit.
        hamiltonians: list of Hamiltonian (one per circuit; QubitOperator accepted).
        params: list of 1-D arrays, item i in ``circuits[i].params_name`` order.
        n_qubits: list of qubit counts (default: each circuit's width).
        dtype: complex64 (default) or complex128.

    Returns:
        f: complex array [n_items] of <psi_i|H_i|psi_i>;
        grads: list of real arrays, grads[i][p] = d f_i / d params[i][p].
    """
    n_items = len(circuits)
    if len(hamiltonians) != n_items or len(params) != n_items:
        raise ValueError("circuits, hamiltonians and params need the same length")
    dt = mqdtype.to_mq_type(dtype)
    code = mqdtype.native_code(dt)
    if n_qubits is None:
        n_qubits = [c.n_qubits for c in circuits]
    if len(n_qubits) != n_items:
        raise ValueError("n_qubits needs one entry per circuit")
    natives, hams, flat, offsets = [], [], [], [0]
    for c, h, p in zip(circuits, hamiltonians, params):
        names = list(c.params_name)
        p = np.asarray(p, dtype=np.float64).reshape(-1)
        if p.size != len(names):
            raise ValueError(f"circuit needs {len(names)} parameters, got {p.size}")
        natives.append(_native_circuit(c, names))
        hams.append(h if isinstance(h, Hamiltonian) else Hamiltonian(h, dt))
        flat.append(p)
        offsets.append(offsets[-1] + p.size)
    flat = np.ascontiguousarray(np.concatenate(flat) if flat else np.zeros(0))
    off = np.asarray(offsets[:-1], dtype=np.int64)
    nq = np.asarray(n_qubits, dtype=np.int32)
    circ_arr = (ctypes.c_void_p * max(n_items, 1))(*[nc.handle.value for nc in natives])
    ham_arr = (ctypes.c_void_p * max(n_items, 1))(*[h.native().value for h in hams])
    f = np.zeros((max(n_items, 1), 2))
    g = np.zeros(max(flat.size, 1))
    _native.check(_native.lib().mqa_grad_items(n_items, circ_arr, ham_arr, _native.ptr(nq), _native.ptr(flat),
                                                _native.ptr(off), code, _native.ptr(f), _native.ptr(g)))
    fc = (f[:n_items, 0] + 1j * f[:n_items, 1])
    grads = [g[offsets[i]:offsets[i + 1]].copy() for i in range(n_items)]
    return fc, grads
