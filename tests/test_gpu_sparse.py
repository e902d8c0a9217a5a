"""GPU parity of sparse (CSR) Hamiltonians: H psi, expectations, the hermitian adjoint gradient and
the non-hermitian gradient (which needs H^dagger) against the dense simulation; a CSR copy of a
Pauli Hamiltonian against the Pauli path and the CPU oracle.  Bars of tests/bars.py against complex128
references, with the largest absolute row sum of the matrix in place of sum |c|."""
import numpy as np
import pytest
import scipy.sparse as sp

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from oracle import cpu, dense
from tests.bars import assert_g, f_tol
from tests.conftest import hea, random_hamiltonian
from tests.test_gate_families import family_circuit

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def tol(dtype, scale=1.0):
    return f_tol(dtype, scale)


def random_csr(n, density, seed, hermitian):
    rng = np.random.default_rng(seed)
    m = sp.random(1 << n, 1 << n, density=density, format="csr", random_state=seed)
    m = m + 1j * sp.random(1 << n, 1 << n, density=density, format="csr", random_state=seed + 1)
    if hermitian:
        m = (m + m.conj().T) * 0.5
    return sp.csr_matrix(m * rng.uniform(0.5, 1.5))


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_and_expectation(dtype):
    n = 8
    m = random_csr(n, 0.05, 3, hermitian=False)
    ham = mq.Hamiltonian(m, dtype)
    circ = family_circuit(n, 30, 4)
    names = list(circ.params_name)
    vals = dict(zip(names, np.random.default_rng(2).uniform(0, 6, size=len(names))))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, vals)
    psi = dense.state(circ, vals, n)
    scale = float(np.abs(m).sum(axis=1).max())
    assert abs(sim.get_expectation(ham) - np.vdot(psi, m @ psi)) < tol(dtype, scale)
    sim.apply_hamiltonian(ham)
    assert np.max(np.abs(sim.get_qs() - m @ psi)) < tol(dtype, scale)


@pytest.mark.parametrize("dtype", DTYPES)
def test_csr_copy_of_pauli_hamiltonian_matches_pauli_path(dtype):
    n, batch = 12, 3
    circ = hea(n, 3, encoder=True)
    op = random_hamiltonian(n, 14, 6)
    h_pauli = mq.Hamiltonian(op, dtype)
    h_csr = mq.Hamiltonian(mq.Hamiltonian(op).sparse_matrix(), dtype)
    rng = np.random.default_rng(7)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f1, ge1, ga1 = sim.get_expectation_with_grad([h_pauli, h_csr], circ)(enc, ans)
    assert _native.get_option("phase_free_used") == _native.get_option("phase_free")  # both hermitian
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    f_ref, g_ref = cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()], enc, ans,
                                  batch_threads=4)
    g1 = np.concatenate([ge1, ga1], axis=2)
    scale = sum(abs(v) for v in op.terms.values())
    for h in range(2):
        assert np.max(np.abs(f1[:, h] - f_ref[:, 0])) < tol(dtype, scale)
        assert_g(g1[:, h], g_ref[:, 0], dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_hermitian_gradient_of_a_general_matrix(dtype):
    """The reference's adjoint with co-state H psi for any matrix H: f = <psi|H|psi> (complex) and
    g = 2 Re <H psi|d psi>."""
    n = 7
    m = random_csr(n, 0.08, 11, hermitian=False)
    circ = family_circuit(n, 25, 12)
    names = list(circ.params_name)
    vals = np.random.default_rng(1).uniform(0, 6, size=len(names))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(m, dtype), circ)(vals)
    assert _native.get_option("phase_free_used") == 0  # <H psi|psi> is not real: exact plans
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, m)
    scale = float(np.abs(m).sum(axis=1).max())
    assert abs(f[0, 0] - f_ref) < tol(dtype, scale)
    assert_g(g[0, 0].real, g_ref, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_nonhermitian_gradient_uses_the_adjoint_matrix(dtype):
    n = 6
    m = random_csr(n, 0.1, 21, hermitian=False)
    cr = family_circuit(n, 20, 22)
    cl = family_circuit(n, 15, 23)
    names = list(cr.params_name) + [k for k in cl.params_name if k not in cr.params_name]
    vals = np.random.default_rng(3).uniform(0, 6, size=len(names))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(m, dtype), cr, cl)(vals)
    f_ref, g_ref = dense.nonhermitian_gradient(cr, cl, dict(zip(names, vals)), names, n, m)
    scale = float(np.abs(m).sum(axis=1).max())
    assert abs(f[0, 0] - f_ref) < tol(dtype, scale)
    assert_g(g[0, 0], g_ref, dtype)


def test_size_must_match():
    sim = mq.Simulator("mqvector_amd", 5)
    with pytest.raises(ValueError):
        sim.get_expectation(mq.Hamiltonian(sp.csr_matrix(np.eye(16)), mq.complex128))
