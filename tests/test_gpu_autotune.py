"""The tile-width autotuner on the GPU: it keeps a candidate, and results are unchanged by the
width it picks (bars of tests/bars.py, here between two runs of the backend)."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, autotune, ir
from mindquantum_amd.algorithm import hardware_efficient_ansatz, layered_random_circuit
from oracle import cpu
from tests.bars import assert_f, assert_g, coeff_sum
from tests.conftest import hea, random_hamiltonian

pytestmark = pytest.mark.gpu


def test_tune_apply_keeps_results():
    n = 18
    circ = layered_random_circuit(n, 8, 11)
    times = autotune.tune_apply(circ, n, mq.complex128, candidates=(9, 10, 11), tolerance=0.0)
    assert set(times) == {9, 10, 11}
    best = min(times, key=times.get)
    assert circ._mqa_tiles[(_native.MQA_C128, False)] == best
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
    sim.apply_circuit(circ)
    ref = cpu.apply_circuit(cpu.zero_state(n), ir.lower(circ, []), np.zeros(0))
    assert 1 - abs(np.vdot(ref, sim.get_qs())) <= 1e-12


@pytest.mark.parametrize("dtype", [mq.complex64, mq.complex128])
def test_tune_grad_keeps_gradients(dtype):
    n, batch = 14, 4
    circ = hea(n, 4, encoder=True)
    op = random_hamiltonian(n, 10, 2)
    rng = np.random.default_rng(3)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)
    f0, ge0, ga0 = grad_ops(enc, ans)
    times = autotune.tune_grad(grad_ops, enc, ans, fwd_candidates=(9, 11), adj_candidates=(9, 10, 11),
                               tolerance=0.0)
    assert len(times) >= 4
    fk, ak = min(times, key=times.get)
    assert grad_ops.native_circuits[0].tile_qubits(mq.dtype.native_code(dtype), True) == ak
    f1, ge1, ga1 = grad_ops(enc, ans)
    assert_f(f1, f0, dtype, coeff_sum(op))
    assert_g(np.concatenate([ge1, ga1], 2), np.concatenate([ge0, ga0], 2), dtype)

