"""GPU parity of the lowered gate families (U3, FSim, Givens, SWAPalpha, Rxy, Rxz, Ryz,
RotPauliString) against the reference's known answer and the dense / CPU oracles.

Bars: the reference test's atol 1e-6 for its known answer; otherwise those of tests/bars.py against
the complex128 oracle (1e-12 fidelity / 1e-10 relative in complex128, 1e-5 relative in complex64).
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import ir
from mindquantum_amd.core import gates as G
from oracle import cpu, dense
from tests.bars import assert_f, assert_g, assert_state, coeff_sum
from tests.test_gate_families import _golden_circuit, family_circuit

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


@pytest.mark.parametrize("dtype", DTYPES)
def test_golden_multi_params_gate_gpu(golden, dtype):
    gv = golden["multi_params_gate"]
    circ = _golden_circuit()
    p0 = np.array(gv["params"])
    sim = mq.Simulator("mqvector_amd", 2, dtype=dtype)
    sim.apply_circuit(circ, pr=dict(zip(circ.params_name, p0)))
    qs_exp = np.array(gv["qs_re"]) + 1j * np.array(gv["qs_im"])
    assert np.allclose(sim.get_qs(), qs_exp, atol=gv["atol"])
    sim.reset()
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator("X0 Y1"), dtype), circ)(p0)
    assert np.allclose(f, gv["f"], atol=gv["atol"])
    assert np.allclose(g, np.array(gv["g"]).reshape(1, 1, 5), atol=gv["atol"])


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("seed", range(2))
def test_family_gradients_gpu(seed, dtype):
    n = 7
    circ = family_circuit(n, 40, 10 + seed)
    op = mq.QubitOperator("Z0 X1") + 0.7 * mq.QubitOperator("Y6") - 0.4 * mq.QubitOperator("X0 Z2 Y3")
    names = list(circ.params_name)
    vals = np.random.default_rng(seed).uniform(0, 2 * np.pi, size=len(names))
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    assert_f(f[0, 0], f_ref, dtype, coeff_sum(op))
    assert_g(g[0, 0].real, g_ref, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_family_state_16q(dtype):
    """Larger register (tile passes with non-local controls and high qubits) against the CPU
    oracle: a 16-qubit circuit of the same families."""
    n = 16
    circ = family_circuit(n, 120, 99)
    names = list(circ.params_name)
    vals = np.random.default_rng(5).uniform(0, 2 * np.pi, size=len(names))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, pr=dict(zip(names, vals)))
    ref = cpu.apply_circuit(cpu.zero_state(n, np.complex128), ir.lower(circ, names), vals)
    assert_state(sim.get_qs(), ref, dtype)
