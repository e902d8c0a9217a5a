"""GPU parity of the compiled (code-generated) tile passes against the CPU oracle.

The options force compilation before the first launch (jit = 2) for every launch size
(jit_min_amps = 0); ``jit_launches`` proves the compiled kernels ran.  Bars of tests/bars.py
against the complex128 oracle.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.core import gates as G
from oracle import dense
from tests.bars import assert_f, coeff_sum
from tests.conftest import hea, params_for, random_circuit, random_hamiltonian
from tests.test_gpu_parity import DTYPES, assert_grad, assert_state, oracle_grad, oracle_state

pytestmark = pytest.mark.gpu


@pytest.fixture
def compiled():
    _native.set_option("jit", 2)
    _native.set_option("jit_min_amps", 0)
    before = _native.get_option("jit_launches")
    yield before
    _native.set_option("jit", -1)
    _native.set_option("jit_min_amps", -1)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_random_circuit_compiled(compiled, dtype):
    n = 13
    circ = random_circuit(n, 150, 9, ctrl_prob=0.3)
    names, vals = params_for(circ, 9)
    values = dict(zip(names, vals))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, values)
    assert_state(sim.get_qs(), oracle_state(circ, n, values, dtype), dtype)
    assert _native.get_option("jit_launches") > compiled


@pytest.mark.parametrize("dtype", DTYPES)
def test_dense_blocks_and_nonlocal_controls_compiled(compiled, dtype):
    n = 15
    rng = np.random.default_rng(3)
    circ = mq.Circuit()
    for q in range(n):
        circ.h(q)
    for k in range(12):
        qs = [int(x) for x in rng.permutation(n)]
        circ += G.RY(float(rng.uniform(0, 6))).on(qs[0], qs[1:3])
        circ += G.X.on(qs[3], [qs[4]])
        m = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
        circ += G.UnivMathGate("u3", m).on(qs[5:8], [qs[8]] if k % 2 else None)
        m4 = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))[0]
        circ += G.UnivMathGate("u4", m4).on(qs[9:13])
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ)
    assert_state(sim.get_qs(), oracle_state(circ, n, {}, dtype), dtype)
    assert _native.get_option("jit_launches") > compiled


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_every_gate_kind_compiled(compiled, dtype):
    n = 7
    circ = random_circuit(n, 80, 77, ctrl_prob=0.3)
    op = random_hamiltonian(n, 10, 5)
    names, vals = params_for(circ, 3)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    assert_f(f[0, 0], f_ref, dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref, dtype)
    assert _native.get_option("jit_launches") > compiled


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_batch_hea_compiled(compiled, dtype):
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
    assert _native.get_option("jit_launches") > compiled
