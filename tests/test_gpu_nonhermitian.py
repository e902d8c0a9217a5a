"""GPU parity of the non-hermitian gradient (circ_left / simulator_left) against the reference's
closed form (test_simulator.py:199-224, 366-382) and the dense product-rule gradient.

Bars (tests/bars.py): the reference test's atol 1e-6 for its closed forms; against the dense
complex128 product rule, 1e-10 (complex128) / 1e-5 (complex64) times max(1, sum |c|) for f and
relative to the largest |g_ref| for g.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from oracle import dense
from tests.bars import assert_f, assert_g, coeff_sum
from tests.test_gate_families import family_circuit

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


@pytest.mark.parametrize("dtype", DTYPES)
def test_closed_form(dtype):
    sim = mq.Simulator("mqvector_amd", 1, dtype=dtype)
    c_r = mq.Circuit().ry("b", 0)
    c_l = mq.Circuit().rz("a", 0)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator(""), dtype), c_r, c_l)
    f, g = grad_ops(np.array([1.2, 2.3]))
    f_exp = np.exp(1j * 2.3 / 2) * np.cos(1.2 / 2)
    g_exp = [-0.5 * np.exp(1j * 2.3 / 2) * np.sin(1.2 / 2), 1j / 2 * np.exp(1j * 2.3 / 2) * np.cos(1.2 / 2)]
    assert f.shape == (1, 1) and g.shape == (1, 1, 2)
    assert np.allclose(f[0, 0], f_exp, atol=1e-6)
    assert np.allclose(g[0, 0], g_exp, atol=1e-6)


@pytest.mark.parametrize("dtype", DTYPES)
def test_fidelity_with_simulator_left(dtype):
    sim1 = mq.Simulator("mqvector_amd", 1, dtype=dtype)
    sim1.apply_circuit(mq.Circuit().h(0))
    sim2 = mq.Simulator("mqvector_amd", 1, dtype=dtype)
    ansatz = mq.Circuit().ry("a", 0).rz("b", 0).ry("c", 0)
    grad_ops = sim2.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator(""), dtype), ansatz, mq.Circuit(),
                                              simulator_left=sim1)
    f, _ = grad_ops(np.array([7.902762e-01, 2.139225e-04, 7.795934e-01]))
    assert np.allclose(np.abs(f), 1, atol=1e-6)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n", [3, 9])
def test_random_pair_against_dense(n, dtype):
    cr = family_circuit(n, 30, 20 + n)
    cl = family_circuit(n, 25, 40 + n)
    names = list(cr.params_name) + [k for k in cl.params_name if k not in cr.params_name]
    vals = np.random.default_rng(n).uniform(0, 2 * np.pi, size=len(names))
    op = mq.QubitOperator("Z0 X1") + 0.5 * mq.QubitOperator("Y2") - 0.25 * mq.QubitOperator(f"X{n - 1}")
    # a prepared left state
    prep = mq.Circuit().h(0).rx(0.7, 1).x(2, 0)
    sim_l = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim_l.apply_circuit(prep)
    init_l = dense.state(prep, {}, n)
    f_ref, g_ref = dense.nonhermitian_gradient(cr, cl, dict(zip(names, vals)), names, n, op, init_l=init_l)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), cr, cl, simulator_left=sim_l)(vals)
    assert list(names) == list(cr.params_name) + [k for k in cl.params_name if k not in cr.params_name]
    assert_f(f[0, 0], f_ref, dtype, coeff_sum(op))
    assert_g(g[0, 0], g_ref, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_encoder_batch_two_hamiltonians(dtype):
    """Encoder rows (batch 3) with the same circuit on both sides (circ_left omitted, a prepared
    simulator_left) and two Hamiltonians: every (sample, hamiltonian) against the dense result."""
    n = 6
    enc = mq.Circuit([mq.RX(f"x{q}").on(q) for q in range(n)]).as_encoder()
    ans = family_circuit(n, 20, 7)
    circ = enc + ans
    sim_l = mq.Simulator("mqvector_amd", n, dtype=dtype)
    prep = mq.Circuit().ry(0.3, 0).h(3)
    sim_l.apply_circuit(prep)
    init_l = dense.state(prep, {}, n)
    ops = [mq.QubitOperator("Z0 Z1"), mq.QubitOperator("X2") + 0.3 * mq.QubitOperator("Y4 Z5")]
    rng = np.random.default_rng(9)
    x = rng.normal(size=(3, n))
    a = rng.uniform(0, 2 * np.pi, size=len(ans.params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, dtype) for o in ops], circ,
                                                    simulator_left=sim_l)(x, a)
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    for b in range(3):
        vals = dict(zip(names, list(x[b]) + list(a)))
        for h, op in enumerate(ops):
            f_ref, g_ref = dense.nonhermitian_gradient(circ, circ, vals, names, n, op, init_l=init_l)
            assert_f(f[b, h], f_ref, dtype, coeff_sum(op))
            assert_g(np.concatenate([g_enc[b, h], g_ans[b, h]]), g_ref, dtype)
