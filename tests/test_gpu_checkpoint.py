"""GPU parity of the checkpointed back-sweep, the large-instance workload shape and the uneven
QAOA batch at full depth, against the CPU oracle.

Checkpointed gradients must equal the plain ones (same kernels, psi taken from a stored forward
state at each back-sweep segment boundary instead of the uncomputed one) and the complex128 oracle
to the bars of tests/bars.py: expectations 1e-10 / 1e-5 times max(1, sum |c|), gradients 1e-10 /
1e-5 relative to the largest |g_ref| (complex128 / complex64).
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.algorithm import hardware_efficient_ansatz, random_pauli_hamiltonian, uneven_qaoa_batch
from oracle import cpu
from tests.bars import assert_f, assert_g, coeff_sum, g_tol
from tests.conftest import hea, random_hamiltonian

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def oracle_grad(circ, n, ops, enc, ans, dtype=None):
    """(f, g) of the complex128 oracle (dtype is ignored)."""
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    return cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(o).terms_ir() for o in ops], enc,
                          ans, batch_threads=4)


def grad_with(options, circ, n, hams, enc, ans, dtype):
    """Gradient with library options set for this call only (enc None: ansatz-only circuit)."""
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    old = {k: _native.get_option(k) for k in options}
    try:
        for k, v in options.items():
            _native.set_option(k, v)
        inputs = (ans,) if enc is None else (enc, ans)
        out = sim.get_expectation_with_grad(hams, circ)(*inputs)
        used = _native.get_option("checkpoints_used")
    finally:
        for k, v in old.items():
            _native.set_option(k, v)
    return out, used


def expected_checkpoints(circ, n, dtype, requested):
    """At most one checkpoint per back-sweep pass boundary."""
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    passes = len(ir.NativeCircuit(circ, names).schedule(n, 0 if dtype == mq.complex64 else 1, adjoint=True))
    return min(requested, passes - 1)


@pytest.mark.parametrize("dtype", DTYPES)
def test_checkpointed_grad_matches_plain_and_oracle(dtype):
    """14 qubits, encoder + HEA-6, batch 6, two Hamiltonians (forward recomputed per Hamiltonian)."""
    n, batch = 14, 6
    circ = hea(n, 6, encoder=True)
    ops = [random_hamiltonian(n, 10, 3), mq.QubitOperator("Z0 Z13") + 0.3 * mq.QubitOperator("Y5 X6")]
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    rng = np.random.default_rng(11)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    (f0, ge0, ga0), used0 = grad_with({"checkpoints": 0}, circ, n, hams, enc, ans, dtype)
    (f1, ge1, ga1), used1 = grad_with({"checkpoints": 3}, circ, n, hams, enc, ans, dtype)
    assert used0 == 0 and used1 == expected_checkpoints(circ, n, dtype, 3) >= 2
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans)
    g0 = np.concatenate([ge0, ga0], axis=2)
    g1 = np.concatenate([ge1, ga1], axis=2)
    for h, o in enumerate(ops):
        assert_f(f1[:, h], f_ref[:, h], dtype, coeff_sum(o))
        assert_g(g1[:, h], g_ref[:, h], dtype)
        assert_g(g0[:, h], g_ref[:, h], dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_checkpoints_limited_by_hbm_budget(dtype):
    """A budget of 3.5 states admits psi, lambda and one checkpoint for one sample: the call
    holds one checkpoint and runs the batch one sample at a time."""
    n, batch = 16, 3
    circ = hea(n, 4, encoder=True)
    op = random_hamiltonian(n, 12, 5)
    rng = np.random.default_rng(12)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    state = (1 << n) * (8 if dtype == mq.complex64 else 16)
    (f, ge, ga), used = grad_with({"checkpoints": 5, "hbm_budget_bytes": int(3.5 * state)}, circ, n,
                                  mq.Hamiltonian(op, dtype), enc, ans, dtype)
    assert used == 1
    f_ref, g_ref = oracle_grad(circ, n, [op], enc, ans)
    assert_f(f, f_ref, dtype, coeff_sum(op))
    assert_g(np.concatenate([ge, ga], axis=2), g_ref, dtype)


def test_checkpoints_bound_c64_uncompute_drift():
    """Deep circuit (12 qubits, HEA-40: 960 parameters) in complex64 against the complex128 oracle.
    With checkpoints every back-sweep segment starts from an exact forward state."""
    n = 12
    circ = hardware_efficient_ansatz(n, 40)
    op = random_pauli_hamiltonian(n, 16, 3)
    rng = np.random.default_rng(13)
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.params_name))
    enc = np.zeros((1, 0))
    _, g_ref = oracle_grad(circ, n, [op], enc, ans)
    errs = {}
    for ck in (0, 8):
        (f, g), used = grad_with({"checkpoints": ck}, circ, n, mq.Hamiltonian(op, mq.complex64), None, ans,
                                 mq.complex64)
        errs[ck] = float(np.max(np.abs(g[0, 0].real - g_ref[0, 0].real)))
    print(f"c64 gradient error vs c128 oracle: no checkpoints {errs[0]:.2e}, 8 checkpoints {errs[8]:.2e}")
    assert errs[8] <= g_tol(mq.complex64, g_ref[0, 0].real)
    assert errs[0] <= g_tol(mq.complex64, g_ref[0, 0].real)


@pytest.mark.parametrize("dtype", DTYPES)
def test_large_instance_shape_20q(dtype):
    """The 32-qubit workload's circuit family at 20 qubits: HEA-4 and 64 random Pauli strings of
    weight 2-4 (flip groups packed into Hamiltonian tile passes), with checkpoints."""
    n = 20
    circ = hardware_efficient_ansatz(n, 4)
    op = random_pauli_hamiltonian(n, 64, 21)
    rng = np.random.default_rng(21)
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.params_name))
    (f, g), used = grad_with({"checkpoints": 4}, circ, n, mq.Hamiltonian(op, dtype), None, ans, dtype)
    assert used == expected_checkpoints(circ, n, dtype, 4) == 4
    f_ref, g_ref = oracle_grad(circ, n, [op], np.zeros((1, 0)), ans)
    assert_f(f[0, 0], f_ref[0, 0], dtype, coeff_sum(op))
    assert_g(g[0, 0].real, g_ref[0, 0].real, dtype)


def test_uneven_qaoa_batch_full_depth():
    """24 circuits of the uneven MaxCut batch (p = 4, qubit counts 12..18, odd counts included)
    in one call, complex64, every item against the oracle."""
    circs, ops, nqs, params = uneven_qaoa_batch(24, 5, n_min=12, n_max=18)
    assert len(set(nqs)) > 3 and any(n % 2 for n in nqs)
    f, grads = mq.expectation_with_grad_batch(circs, [mq.Hamiltonian(o, mq.complex64) for o in ops], params, nqs,
                                              mq.complex64)
    for i, (c, o, n, p) in enumerate(zip(circs, ops, nqs, params)):
        f_ref, g_ref = cpu.grad_batch(n, mq.complex128, ir.lower(c, list(c.params_name)),
                                      [mq.Hamiltonian(o).terms_ir()], np.zeros((1, 0)), p, batch_threads=1)
        assert_f(f[i], f_ref[0, 0], mq.complex64, coeff_sum(o), f"f of item {i} (n={n})")
        assert_g(grads[i], g_ref[0, 0].real, mq.complex64, f"g of item {i} (n={n})", csum=coeff_sum(o, identity=False))
