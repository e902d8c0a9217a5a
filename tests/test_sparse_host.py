"""Sparse (CSR) Hamiltonians on the host: construction rules of hamiltonian.py:56-76, the
Pauli-to-sparse conversion against the dense Kronecker form, and C ABI validation of the CSR
arrays (no GPU needed)."""
import ctypes

import numpy as np
import pytest
import scipy.sparse as sp

import mindquantum_amd as mq
from mindquantum_amd import _native
from oracle import dense
from tests.conftest import random_hamiltonian


def test_pauli_to_sparse_matches_dense():
    op = random_hamiltonian(5, 12, 3)
    h = mq.Hamiltonian(op)
    assert np.allclose(h.sparse_matrix().toarray(), dense.hamiltonian_matrix(op, 5), atol=1e-14)
    assert h.sparse(6).how_to == 1 and h.n_qubits == 6
    with pytest.raises(ValueError):
        h.sparse(7)


def test_csr_construction_rules():
    m = sp.random(16, 16, density=0.2, format="csr", random_state=1) * (1 + 0.5j)
    h = mq.Hamiltonian(m, mq.complex64)
    assert h.how_to == 2 and h.n_qubits == 4
    assert np.allclose(h.sparse_matrix().toarray(), m.toarray())
    with pytest.raises(ValueError):
        mq.Hamiltonian(sp.csr_matrix(np.eye(6)))
    with pytest.raises(ValueError):
        mq.Hamiltonian(sp.csr_matrix(np.ones((4, 8))))
    with pytest.raises(TypeError):
        mq.Hamiltonian(sp.coo_matrix(np.eye(4)))
    with pytest.raises(ValueError):
        h.terms_ir()
    assert h.native().value


def test_csr_abi_rejects_bad_arrays():
    out = ctypes.c_void_p()
    ptr = np.array([0, 1, 2], dtype=np.int64)
    idx = np.array([0, 5], dtype=np.int64)  # column 5 of a 2 x 2 matrix
    val = np.ones(4)
    st = _native.lib().mqa_ham_create_csr(2, 2, _native.ptr(ptr), _native.ptr(idx), _native.ptr(val), ctypes.byref(out))
    assert st == _native.MQA_ERR_ARG
    st = _native.lib().mqa_ham_create_csr(3, 2, _native.ptr(ptr), _native.ptr(idx), _native.ptr(val), ctypes.byref(out))
    assert st == _native.MQA_ERR_ARG
