"""Gradient passes that read the initial state directly and skip dead outputs: the first forward
pass loads every sample's tiles from the simulator's state (no copy into the workspace first), and
the last back-sweep pass writes only its taps.  Checked from a prepared (non-|0>) state, batched,
with several Hamiltonians, checkpoints, interpreted and compiled passes, against the complex128
oracle; the simulator's own state is left unchanged."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from oracle import dense
from tests.bars import assert_f, assert_g, coeff_sum
from tests.conftest import hea, random_circuit, random_hamiltonian
from tests.test_gpu_parity import oracle_grad

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


@pytest.fixture
def opts():
    keys = ("jit", "jit_min_amps", "checkpoints")
    saved = {k: _native.get_option(k) for k in keys}
    yield
    for k, v in saved.items():
        _native.set_option(k, v)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compiled", [False, True])
@pytest.mark.parametrize("checkpoints", [0, 2])
def test_prepared_state_batch(opts, dtype, compiled, checkpoints):
    n, batch = 13, 5
    _native.set_option("jit", 2 if compiled else 0)
    _native.set_option("jit_min_amps", 0)
    _native.set_option("checkpoints", checkpoints)
    prep = random_circuit(n, 40, 8, param_names=False)
    circ = hea(n, 2, encoder=True)
    ops = [random_hamiltonian(n, 6, 3), mq.QubitOperator("Z0 Z12") + 0.4 * mq.QubitOperator("X6")]
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(prep)
    init = sim.get_qs()
    rng = np.random.default_rng(13)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, dtype) for o in ops], circ)(enc, ans)
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans, init=init)
    for h, o in enumerate(ops):
        assert_f(f[:, h], f_ref[:, h], dtype, coeff_sum(o))
    assert_g(np.concatenate([g_enc, g_ans], axis=2), g_ref, dtype)
    np.testing.assert_array_equal(sim.get_qs(), init)  # the gradient works on copies


@pytest.mark.parametrize("dtype", DTYPES)
def test_gate_free_circuit_still_reads_the_state(dtype):
    """A circuit without gates has no passes: f = <psi|H|psi> of the simulator's state."""
    n = 7
    prep = random_circuit(n, 20, 2, param_names=False)
    op = random_hamiltonian(n, 5, 9)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(prep)
    init = sim.get_qs().astype(np.complex128)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), mq.Circuit())
    f, g = grad_ops(np.zeros(0))
    ref = np.vdot(init, dense.hamiltonian_matrix(op, n) @ init)
    assert_f(f[0, 0], ref, dtype, coeff_sum(op))
    assert g.shape[-1] == 0
