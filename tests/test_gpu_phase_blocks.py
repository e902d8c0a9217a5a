"""GPU parity of phase blocks (option "phase_ops") against the complex128 oracle.

Two executions of the same plans: compiled passes (phase_ops 1 with jit = 2 and jit_min_amps = 0,
so every launch compiles) and the interpreted pass (phase_ops 2 with jit = 0).  Circuits are rich in
RZ / PS / RZZ gates, controlled and not, next to the other kinds; gradients run both the
phase-free plans (hermitian observables, the default) and the exact ones (option phase_free 0).
Bars of tests/bars.py.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.algorithm.builders import max_cut_ansatz, max_cut_hamiltonian, qaoa_graph
from mindquantum_amd.core import gates as G
from oracle import dense
from tests.bars import assert_f, coeff_sum
from tests.conftest import hea, params_for, random_hamiltonian
from tests.test_gpu_parity import DTYPES, assert_grad, assert_state, oracle_grad, oracle_state
from tests.test_phase_blocks import phase_circuit

pytestmark = pytest.mark.gpu

MODES = ["compiled", "interpreted"]


@pytest.fixture(params=MODES)
def phase_mode(request):
    if request.param == "compiled":
        _native.set_option("phase_ops", 1)
        _native.set_option("jit", 2)
        _native.set_option("jit_min_amps", 0)
    else:
        _native.set_option("phase_ops", 2)
        _native.set_option("jit", 0)
    before = _native.get_option("jit_launches")
    yield request.param, before
    for k in ("phase_ops", "jit", "jit_min_amps"):
        _native.set_option(k, -1)


def _check_launches(phase_mode):
    mode, before = phase_mode
    if mode == "compiled":
        assert _native.get_option("jit_launches") > before
    else:
        assert _native.get_option("jit_launches") == before


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,seed", [(9, 1), (14, 2)])
def test_apply_phase_circuit(phase_mode, dtype, n, seed):
    circ = phase_circuit(n, 180, seed)
    names, vals = params_for(circ, seed)
    values = dict(zip(names, vals))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, values)
    assert_state(sim.get_qs(), oracle_state(circ, n, values, dtype), dtype)
    _check_launches(phase_mode)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("phase_free", [1, 0])
def test_grad_phase_circuit(phase_mode, dtype, phase_free):
    """Every kind next to controlled and uncontrolled RZ / PS / RZZ; the dense product rule."""
    n = 8
    circ = phase_circuit(n, 90, 21)
    op = random_hamiltonian(n, 10, 6)
    names, vals = params_for(circ, 4)
    _native.set_option("phase_free", phase_free)
    try:
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    finally:
        _native.set_option("phase_free", -1)
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    assert_f(f[0, 0], f_ref, dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref, dtype)
    _check_launches(phase_mode)


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_batch_hea(phase_mode, dtype):
    """Encoder + HEA batch on 16 qubits: RZ layers as blocks over register, thread and tile bits."""
    n, batch = 16, 4
    circ = hea(n, 3, encoder=True)
    ops = [mq.QubitOperator("Z0 Z1") + 0.5 * mq.QubitOperator("X15"), random_hamiltonian(n, 8, 16)]
    rng = np.random.default_rng(16)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, dtype) for o in ops], circ)(enc, ans)
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans)
    for h, o in enumerate(ops):
        assert_f(f[:, h], f_ref[:, h], dtype, coeff_sum(o))
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, dtype)
    _check_launches(phase_mode)


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_qaoa(phase_mode, dtype):
    """MaxCut QAOA: RZZ layers on every edge as parity blocks, with checkpoints."""
    n = 14
    graph = qaoa_graph(n, 9)
    circ = max_cut_ansatz(graph, 3)
    op = max_cut_hamiltonian(graph)
    rng = np.random.default_rng(9)
    ans = rng.uniform(0, np.pi, size=len(circ.ansatz_params_name))
    _native.set_option("checkpoints", 2)
    try:
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(ans)
    finally:
        _native.set_option("checkpoints", -1)
    f_ref, g_ref = oracle_grad(circ, n, [op], np.zeros((1, 0)), ans)
    assert_f(f[0, 0], f_ref[0, 0], dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref[0, 0], dtype)
    _check_launches(phase_mode)


@pytest.mark.parametrize("dtype", DTYPES)
def test_controlled_phases_on_nonlocal_qubits(phase_mode, dtype):
    """QFT-like controlled phases whose qubits span the tile bits of an 18-qubit state."""
    n = 18
    circ = mq.Circuit()
    for q in range(n):
        circ.h(q)
    rng = np.random.default_rng(18)
    for k in range(60):
        a, b, c = (int(x) for x in rng.permutation(n)[:3])
        circ += G.PhaseShift(f"p{k}").on(a, [b])
        circ += G.RZ(f"r{k}").on(b, [c] if k % 3 == 0 else None)
        circ += G.Rzz(f"z{k}").on([a, c])
        if k % 5 == 0:
            circ += G.RY(f"y{k}").on(c)
    names, vals = params_for(circ, 18)
    values = dict(zip(names, vals))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, values)
    assert_state(sim.get_qs(), oracle_state(circ, n, values, dtype), dtype)
    _check_launches(phase_mode)
