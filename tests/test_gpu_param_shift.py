"""Parameter-shift gradients (pr_shift=True) on the GPU against the adjoint gradient, after the
reference's tests test_method_of_mqvector.py:195-285: one- and two-parameter rotations with
two-term parameter expressions and U3 (exact shift rule; reference bars 1e-4 / 1e-3, plus the bars of
tests/bars.py against the complex128 adjoint gradient, itself pinned to the complex128 oracle), and
controlled gates and FSim (central differences with a 1e-3 step; reference bars 1e-3 / 1e-2: in
complex64 a difference quotient of two float expectations carries about 1e-4 of rounding, so only f is
held to the 1e-5 bar there)."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd.core import gates as G
from tests.bars import assert_f, assert_g
from tests.conftest import random_circuit

pytestmark = pytest.mark.gpu

SINGLE_PARAMETER = [G.RX, G.RY, G.RZ, G.PhaseShift, G.Rxx, G.Ryy, G.Rzz, G.Rxy, G.Rxz, G.Ryz, G.SWAPalpha,
                    G.GlobalPhase, G.Givens]
FILLER = ["X", "Y", "Z", "H", "S", "T", "SX", "SWAP", "ISWAP", "CNOT"]


def shift_vs_adjoint(circ, dtype, seed):
    """(f, g) by the parameter-shift rule in dtype, and the adjoint (f, g) in complex128."""
    op = mq.QubitOperator("X0") + mq.QubitOperator("Z0")
    sim = mq.Simulator("mqvector_amd", 3, dtype=dtype)
    pr = np.random.default_rng(seed).random(len(circ.params_name)) * 2 * np.pi
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ, pr_shift=True)(pr)
    sim128 = mq.Simulator("mqvector_amd", 3, dtype=mq.complex128)
    ref_f, ref_g = sim128.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex128), circ)(pr)
    assert f.shape == ref_f.shape and g.shape == ref_g.shape
    return f, g, ref_f, ref_g


@pytest.mark.parametrize("dtype", [mq.complex64, mq.complex128])
def test_parameter_shift_rule(dtype):
    circ = random_circuit(3, 10, 1, kinds=FILLER, ctrl_prob=0.0)
    for i, gate in enumerate(SINGLE_PARAMETER):
        gate_ = gate({f"pr0_{i}": 1, f"pr1_{i}": 2})
        if isinstance(gate_, G.Givens):
            continue  # central differences (not an exact rule): covered below
        circ += gate_.on(list(range(gate_.n_qubits)))
        circ += random_circuit(3, 10, 10 + i, kinds=FILLER, ctrl_prob=0.0)
    circ += G.U3("u3_theta", "u3_phi", "u3_lamda").on(0)
    circ += random_circuit(3, 10, 50, kinds=FILLER, ctrl_prob=0.0)
    circ += G.U3("u3_theta_1", "u3_phi_1", 1).on(0)
    f, g, ref_f, ref_g = shift_vs_adjoint(circ, dtype, 7)
    assert np.allclose(f, ref_f, atol=1e-4)
    assert np.allclose(g, ref_g, atol=1e-3)
    assert_f(f, ref_f, dtype, 2.0)
    if dtype == mq.complex128:
        assert np.max(np.abs(g - ref_g)) < 1e-8
    else:
        assert_g(g, ref_g, dtype)


@pytest.mark.parametrize("dtype", [mq.complex64, mq.complex128])
def test_parameter_shift_rule_finite_diff_case(dtype):
    circ = random_circuit(3, 10, 2, kinds=FILLER, ctrl_prob=0.0)
    for i, gate in enumerate(SINGLE_PARAMETER):
        gate_ = gate({f"pr0_{i}": 1, f"pr1_{i}": 2})
        if gate_.n_qubits > 2:
            continue
        circ += gate_.on(list(range(gate_.n_qubits)), 2)
        circ += random_circuit(3, 10, 20 + i, kinds=FILLER, ctrl_prob=0.0)
    circ += G.U3("u3_theta", "u3_phi", "u3_lamda").on(0, 1)
    circ += random_circuit(3, 10, 60, kinds=FILLER, ctrl_prob=0.0)
    circ += G.FSim("fsim_theta", "fsim_phi").on([0, 1])
    f, g, ref_f, ref_g = shift_vs_adjoint(circ, dtype, 8)
    assert np.allclose(f, ref_f, atol=1e-3)
    assert np.allclose(g, ref_g, atol=1e-2)
    assert_f(f, ref_f, dtype, 2.0)
    if dtype == mq.complex128:
        assert np.max(np.abs(g - ref_g)) < 1e-5  # central differences with a 1e-3 step


def test_parameter_shift_with_encoder_batch():
    n = 4
    enc = mq.Circuit([G.RX(f"x{q}").on(q) for q in range(n)]).as_encoder()
    ans = mq.Circuit([G.RY(f"a{q}").on(q) for q in range(n)] + [G.X.on(q + 1, q) for q in range(n - 1)] +
                     [G.U3(f"t{q}", f"p{q}", 0.4).on(q) for q in range(n)])
    circ = enc + ans
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
    hams = [mq.Hamiltonian(mq.QubitOperator("Z0 Z1"), mq.complex128),
            mq.Hamiltonian(mq.QubitOperator("X2") + 0.5 * mq.QubitOperator("Y3"), mq.complex128)]
    rng = np.random.default_rng(4)
    x = rng.normal(size=(3, n))
    a = rng.uniform(0, 6, size=len(ans.params_name))
    f, ge, ga = sim.get_expectation_with_grad(hams, circ, pr_shift=True)(x, a)
    rf, rge, rga = sim.get_expectation_with_grad(hams, circ)(x, a)
    assert np.max(np.abs(f - rf)) < 1e-10
    assert np.max(np.abs(ge - rge)) < 1e-8 and np.max(np.abs(ga - rga)) < 1e-8
