"""GPU parity of the uneven circuit batch (mqa_grad_items) against the CPU oracle.

Items of different qubit counts, schedules and Hamiltonians run through shared launches; every
item must match the oracle's per-circuit gradient (the reference's one-call-per-circuit result).
Bars of tests/bars.py against the complex128 oracle.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.algorithm import max_cut_ansatz, max_cut_hamiltonian, random_regular_graph
from oracle import cpu
from tests.bars import assert_f, assert_g, coeff_sum
from tests.conftest import hea, random_circuit, random_hamiltonian

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def oracle_item(circ, op, n, params):
    """(f, g) of one item from the complex128 oracle."""
    names = list(circ.params_name)
    f, g = cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()], np.zeros((1, 0)),
                          np.asarray(params, dtype=np.float64), batch_threads=1)
    return f[0, 0], g[0, 0].real


def check_items(circs, ops, nqs, params, dtype):
    f, grads = mq.expectation_with_grad_batch(circs, [mq.Hamiltonian(o, dtype) for o in ops], params, nqs, dtype)
    assert f.shape == (len(circs),) and len(grads) == len(circs)
    for i, (c, o, n, p) in enumerate(zip(circs, ops, nqs, params)):
        f_ref, g_ref = oracle_item(c, o, n, p)
        assert_f(f[i], f_ref, dtype, coeff_sum(o), f"f of item {i} (n={n})")
        assert grads[i].shape == g_ref.shape
        if g_ref.size:
            assert_g(grads[i], g_ref, dtype, f"g of item {i} (n={n})", csum=coeff_sum(o, identity=False))


@pytest.mark.parametrize("dtype", DTYPES)
def test_uneven_qaoa_batch(dtype):
    """MaxCut QAOA (p = 2) on random 3-regular graphs of 6 to 16 qubits in one call."""
    rng = np.random.default_rng(4)
    circs, ops, nqs, params = [], [], [], []
    for i, n in enumerate([6, 8, 12, 14, 16, 10, 12, 16, 8]):
        graph = random_regular_graph(n, 3, 100 + i)
        c = max_cut_ansatz(graph, 2)
        circs.append(c)
        ops.append(max_cut_hamiltonian(graph))
        nqs.append(n)
        params.append(rng.uniform(0, np.pi, size=len(c.params_name)))
    check_items(circs, ops, nqs, params, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_mixed_items(dtype):
    """Different circuit families, a qubit count above the circuit width, every gate kind, a
    Hamiltonian whose flip mask fits no tile (gather kernel) and a circuit without parameters."""
    rng = np.random.default_rng(9)
    circs, ops, nqs, params = [], [], [], []
    c = hea(13, 2)
    circs.append(c), ops.append(random_hamiltonian(13, 12, 1)), nqs.append(15)
    c = random_circuit(9, 70, 3, ctrl_prob=0.3)
    circs.append(c), ops.append(random_hamiltonian(9, 8, 2)), nqs.append(9)
    c = hea(20, 1)
    circs.append(c), ops.append(mq.QubitOperator(" ".join(f"X{q}" for q in range(20)), 0.4)
                                + mq.QubitOperator("Z0 Z19", 0.3)), nqs.append(20)
    c = mq.Circuit().h(0).x(1, 0).h(2)
    circs.append(c), ops.append(mq.QubitOperator("Z0 Z1") + mq.QubitOperator("X2", 0.5)), nqs.append(3)
    for c in circs:
        params.append(rng.uniform(0, 2 * np.pi, size=len(c.params_name)))
    check_items(circs, ops, nqs, params, dtype)


def test_items_in_budget_chunks():
    """A budget of about one large state: the items run in several chunks with the same result."""
    dtype = mq.complex64
    rng = np.random.default_rng(2)
    circs, ops, nqs, params = [], [], [], []
    for i, n in enumerate([16, 14, 16, 12, 10]):
        graph = random_regular_graph(n, 3, 7 + i)
        c = max_cut_ansatz(graph, 1)
        circs.append(c), ops.append(max_cut_hamiltonian(graph)), nqs.append(n)
        params.append(rng.uniform(0, np.pi, size=len(c.params_name)))
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    f0, g0 = mq.expectation_with_grad_batch(circs, hams, params, nqs, dtype)
    _native.set_option("hbm_budget_bytes", 3 * (1 << 16) * 8)
    try:
        f1, g1 = mq.expectation_with_grad_batch(circs, hams, params, nqs, dtype)
    finally:
        _native.set_option("hbm_budget_bytes", -1)
    assert np.max(np.abs(f1 - f0)) < 1e-6
    for a, b in zip(g0, g1):
        assert np.max(np.abs(a - b)) < 1e-6
    check_items(circs, ops, nqs, params, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_solo_items_on_compiled_passes(dtype):
    """Items at or above the "items_jit_qubits" size run one by one on the per-circuit compiled
    passes (the batched gradient of one sample), the rest in the shared launches: both kinds in one
    call match the oracle and the all-shared call (identity parts, a qubit count above the circuit
    width, MaxCut Hamiltonians with their constant)."""
    rng = np.random.default_rng(12)
    circs, ops, nqs, params = [], [], [], []
    for i, n in enumerate([12, 8, 14, 10, 16]):
        graph = random_regular_graph(n, 3, 40 + i)
        c = max_cut_ansatz(graph, 2)
        circs.append(c), ops.append(max_cut_hamiltonian(graph)), nqs.append(n)
    c = hea(11, 2)
    circs.append(c), ops.append(random_hamiltonian(11, 10, 5) + mq.QubitOperator("", 0.25)), nqs.append(13)
    for c in circs:
        params.append(rng.uniform(0, 2 * np.pi, size=len(c.params_name)))
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    jit0 = _native.get_option("jit")
    _native.set_option("items_jit_qubits", 0)
    try:
        f0, g0 = mq.expectation_with_grad_batch(circs, hams, params, nqs, dtype)
        _native.set_option("jit", 2)
        _native.set_option("items_jit_qubits", 12)
        launches = _native.get_option("jit_launches")
        f1, g1 = mq.expectation_with_grad_batch(circs, hams, params, nqs, dtype)
        assert _native.get_option("jit_launches") > launches  # the solo items ran compiled passes
    finally:
        _native.set_option("items_jit_qubits", -1)
        _native.set_option("jit", jit0)
    tol = 1e-5 if dtype == mq.complex64 else 1e-11
    scale = max(1.0, float(np.max(np.abs(f0))))
    assert np.max(np.abs(f1 - f0)) < tol * scale
    for a, b in zip(g0, g1):
        assert np.max(np.abs(a - b)) < tol * max(1.0, float(np.max(np.abs(b))))
    check_items(circs, ops, nqs, params, dtype)
