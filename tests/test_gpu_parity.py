"""GPU parity of the HIP backend (through the C ABI) against the complex128 CPU oracle.

Bars (tests/bars.py): complex128 states - fidelity 1 - |<ref|psi>| <= 1e-12; complex64 - max
amplitude error <= 1e-5 (unit norm); expectations 1e-10 (complex128) / 1e-5 (complex64) times
max(1, sum |c|); gradients 1e-10 / 1e-5 relative to the largest |g_ref|.  Every reference value is
computed in complex128, whatever the precision of the GPU run.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.algorithm import (hardware_efficient_ansatz, layered_random_circuit, max_cut_ansatz,
                                       max_cut_hamiltonian, random_pauli_hamiltonian, random_regular_graph)
from mindquantum_amd.algorithm.builders import encoder_rx
from mindquantum_amd.core import gates as G
from oracle import cpu, dense
from tests.bars import assert_f, assert_g, assert_state, coeff_sum
from tests.conftest import hea, params_for, random_circuit, random_hamiltonian

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def oracle_state(circ, n, values, dtype=None, init=None):
    """State of the complex128 oracle (dtype is ignored: references are always complex128)."""
    names = list(circ.params_name)
    qs = cpu.zero_state(n, np.complex128) if init is None else np.array(init, dtype=np.complex128)
    return cpu.apply_circuit(qs, ir.lower(circ, names), np.array([values[k] for k in names], dtype=np.float64))


def oracle_grad(circ, n, ops, enc, ans, dtype=None, init=None):
    """(f, g) of the complex128 oracle (dtype is ignored: references are always complex128)."""
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    init = None if init is None else np.asarray(init, dtype=np.complex128)
    return cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(o).terms_ir() for o in ops], enc,
                          ans, batch_threads=4, init=init)


def assert_grad(g, g_ref, dtype):
    assert_g(g, g_ref, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,seed", [(1, 0), (3, 1), (7, 2), (10, 3), (13, 4), (17, 5)])
def test_apply_circuit_random(n, seed, dtype):
    circ = random_circuit(n, 120, seed)
    names, vals = params_for(circ, seed)
    values = dict(zip(names, vals))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ, values)
    assert_state(sim.get_qs(), oracle_state(circ, n, values, dtype), dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_circuit_20q_layers(dtype):
    n = 20
    circ = layered_random_circuit(n, 6, 3)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ)
    assert_state(sim.get_qs(), oracle_state(circ, n, {}, dtype), dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_non_local_controls_and_dense_blocks(dtype):
    """Controls on qubits outside the tile, 3- and 4-qubit matrix gates, every tile size."""
    n = 16
    rng = np.random.default_rng(7)
    circ = mq.Circuit()
    for q in range(n):
        circ.h(q)
    for k in range(30):
        qs = [int(x) for x in rng.permutation(n)]
        circ += G.RY(float(rng.uniform(0, 6))).on(qs[0], qs[1:3])
        circ += G.X.on(qs[3], [qs[4]])
        m = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
        circ += G.UnivMathGate("u3", m).on(qs[5:8], [qs[8]] if k % 2 else None)
        m4 = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))[0]
        circ += G.UnivMathGate("u4", m4).on(qs[9:13])
    ref = oracle_state(circ, n, {}, dtype)
    key = "tile_qubits_c64" if dtype == mq.complex64 else "tile_qubits_c128"
    for k in (8, 10, 12):
        _native.set_option(key, k)
        try:
            sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
            sim.apply_circuit(circ)
            assert_state(sim.get_qs(), ref, dtype)
        finally:
            _native.set_option(key, -1)


@pytest.mark.parametrize("dtype", DTYPES)
def test_state_io_reset_copy_astype(dtype):
    n = 9
    rng = np.random.default_rng(1)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.set_qs(psi)
    assert np.allclose(sim.get_qs(), psi / np.linalg.norm(psi), atol=1e-6)
    c = sim.copy()
    sim.reset()
    z = np.zeros(1 << n)
    z[0] = 1
    assert np.allclose(sim.get_qs(), z)
    assert np.allclose(c.get_qs(), psi / np.linalg.norm(psi), atol=1e-6)
    other = mq.complex128 if dtype == mq.complex64 else mq.complex64
    assert np.allclose(c.astype(other).get_qs(), c.get_qs(), atol=1e-6)
    with pytest.raises(ValueError):
        sim.set_qs(np.ones(3))


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,seed", [(2, 0), (6, 1), (12, 2), (18, 3)])
def test_expectation(n, seed, dtype):
    circ = random_circuit(n, 60, 10 + seed)
    names, vals = params_for(circ, seed)
    values = dict(zip(names, vals))
    op = random_hamiltonian(n, 15, seed)
    ham = mq.Hamiltonian(op, dtype)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    got = sim.get_expectation(ham, circ, pr=values)
    q = oracle_state(circ, n, values)
    ref = cpu.expectation(q, q, mq.Hamiltonian(op, mq.complex128).terms_ir())
    assert_f(got, ref, dtype, coeff_sum(op))
    # the simulator's own state is untouched
    z = np.zeros(1 << n)
    z[0] = 1
    assert np.allclose(sim.get_qs(), z)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_hamiltonian_and_lr_expectation(dtype):
    n = 8
    op = random_hamiltonian(n, 10, 4)
    ham = mq.Hamiltonian(op, dtype)
    circ = random_circuit(n, 40, 21, param_names=False)
    circ_l = random_circuit(n, 40, 22, param_names=False)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ)
    psi = sim.get_qs().astype(np.complex128)
    left = mq.Simulator("mqvector_amd", n, dtype=dtype)
    left.apply_circuit(circ_l)
    phi = left.get_qs().astype(np.complex128)
    hm = dense.hamiltonian_matrix(op, n)
    val = sim.get_expectation(ham, mq.Circuit(), mq.Circuit(), left)
    csum = coeff_sum(op)
    assert_f(val, np.vdot(phi, hm @ psi), dtype, csum)
    sim.apply_hamiltonian(ham)
    assert_f(sim.get_qs(), hm @ psi, dtype, csum, "H psi")
    ip = mq.inner_product(left, left)
    assert_f(ip, 1.0, dtype)


def test_golden_values_through_the_gpu(golden):
    for dtype in DTYPES:
        gv = golden["ry_z0_docstring"]
        sim = mq.Simulator("mqvector_amd", 1, dtype=dtype)
        grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator("Z0"), dtype),
                                                 mq.Circuit().ry("a", 0))
        f, g = grad_ops(np.array([gv["a"]]))
        assert f.shape == (1, 1) and g.shape == (1, 1, 1)
        assert np.allclose(f, gv["f"], atol=1e-6) and np.allclose(g, gv["g"], atol=1e-6)
        gv = golden["mul_qubit_gate"]
        x = gv["x"]
        diag = np.diag([np.exp(1j * 2 * (k + 1) * x) for k in range(8)])
        circ = mq.UN(G.H, 3) + G.UnivMathGate("q", diag).on([0, 1, 2])
        sim = mq.Simulator("mqvector_amd", 3, dtype=dtype)
        assert np.allclose(sim.get_expectation(mq.Hamiltonian(mq.QubitOperator("X0"), dtype), circ), gv["f"],
                           atol=1e-6)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,batch", [(4, 3), (12, 8), (16, 2)])
def test_grad_batch_encoder_ansatz(n, batch, dtype):
    circ = hea(n, 3, encoder=True)
    ops = [mq.QubitOperator("Z0 Z1") + 0.5 * mq.QubitOperator(f"X{n - 1}"), random_hamiltonian(n, 8, n)]
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    rng = np.random.default_rng(n)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g_enc, g_ans = sim.get_expectation_with_grad(hams, circ)(enc, ans)
    assert f.shape == (batch, 2) and g_enc.shape == (batch, 2, n) and g_ans.shape == (batch, 2, len(ans))
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans)
    for h, o in enumerate(ops):
        assert_f(f[:, h], f_ref[:, h], dtype, coeff_sum(o))
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_random_circuits_every_gate_kind(dtype):
    n = 7
    circ = random_circuit(n, 80, 77, ctrl_prob=0.3)
    op = random_hamiltonian(n, 10, 5)
    names, vals = params_for(circ, 3)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    assert_f(f[0, 0], f_ref, dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref, dtype)


def test_grad_central_fd_c128():
    """Adjoint vs central finite differences, h = 1e-3 in complex128."""
    n = 10
    circ = hea(n, 2, encoder=True)
    op = random_hamiltonian(n, 12, 9)
    ham = mq.Hamiltonian(op, mq.complex128)
    rng = np.random.default_rng(2)
    enc = rng.normal(size=(1, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
    grad_ops = sim.get_expectation_with_grad(ham, circ)
    _, g_enc, g_ans = grad_ops(enc, ans)
    h = 1e-3
    for k in range(0, len(ans), 7):
        ap, am = ans.copy(), ans.copy()
        ap[k] += h
        am[k] -= h
        fd = (grad_ops(enc, ap)[0] - grad_ops(enc, am)[0]).real / (2 * h)
        assert abs(fd[0, 0] - g_ans[0, 0, k].real) < 1e-5


def test_grad_chunking_and_determinism():
    n, batch = 10, 37
    circ = hea(n, 2, encoder=True)
    ham = mq.Hamiltonian(random_hamiltonian(n, 6, 1), mq.complex64)
    rng = np.random.default_rng(5)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
    ops = sim.get_expectation_with_grad(ham, circ)
    f1, ge1, ga1 = ops(enc, ans)
    f2, ge2, ga2 = ops(enc, ans)
    assert np.array_equal(f1, f2) and np.array_equal(ge1, ge2) and np.array_equal(ga1, ga2)
    _native.set_option("hbm_budget_bytes", 5 * (1 << n) * 8 * 3)  # a few samples per chunk
    try:
        f3, ge3, ga3 = ops(enc, ans)
    finally:
        _native.set_option("hbm_budget_bytes", -1)
    assert np.allclose(f1, f3, atol=1e-6) and np.allclose(ga1, ga3, atol=1e-6)


def test_grad_from_prepared_state_and_no_grad():
    n = 6
    prep = random_circuit(n, 30, 3, param_names=False)
    circ = hea(n, 2)
    circ[0].coeff.no_grad()
    op = random_hamiltonian(n, 5, 2)
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
    sim.apply_circuit(prep)
    init = sim.get_qs()
    names, vals = params_for(circ, 1)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex128), circ)(vals)
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op, init=init)
    assert abs(f[0, 0] - f_ref) < 1e-10
    assert np.allclose(g[0, 0].real, g_ref, atol=1e-9)
    assert g[0, 0, 0] == 0


def test_qaoa_maxcut_small():
    graph = random_regular_graph(12, 3, 4)
    circ = max_cut_ansatz(graph, 4)
    op = max_cut_hamiltonian(graph)
    rng = np.random.default_rng(0)
    vals = rng.uniform(0, np.pi, size=len(circ.params_name))
    for dtype in DTYPES:
        sim = mq.Simulator("mqvector_amd", 12, dtype=dtype)
        f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
        f_ref, g_ref = oracle_grad(circ, 12, [op], np.zeros((1, 0)), vals)
        assert_f(f[0, 0], f_ref[0, 0], dtype, coeff_sum(op))
        assert_grad(g, g_ref, dtype)


def test_hea16_c128_workload():
    """The CPU-runnable workload: 16-qubit HEA in complex128, 10 layers of RY, RZ on every qubit and
    a CNOT ladder (320 parameters ~ U[0, 2pi) from seed 0), H = sum_i Z_i Z_{i+1} (open chain) +
    sum_i 0.5 X_i; one parameter vector.  f and all 320 gradients against the oracle."""
    n = 16
    circ = hardware_efficient_ansatz(n, 10)
    op = mq.QubitOperator("", 0)
    for i in range(n - 1):
        op += mq.QubitOperator(f"Z{i} Z{i + 1}")
    for i in range(n):
        op += 0.5 * mq.QubitOperator(f"X{i}")
    vals = np.random.default_rng(0).uniform(0, 2 * np.pi, size=len(circ.params_name))
    assert len(vals) == 320
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex128), circ)(vals)
    f_ref, g_ref = oracle_grad(circ, n, [op], np.zeros((1, 0)), vals)
    assert_f(f[0, 0], f_ref[0, 0], mq.complex128, coeff_sum(op))
    assert_grad(g, g_ref, mq.complex128)


def test_torch_autograd_function():
    import torch

    from mindquantum_amd.framework import MQAnsatzOnlyOps, MQLayer

    circ = mq.Circuit().rx("a", 0)
    sim = mq.Simulator("mqvector_amd", 1, dtype=mq.complex64)
    ops = MQAnsatzOnlyOps(sim.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator("Z0"), mq.complex64), circ))
    a = torch.tensor([0.5], requires_grad=True)
    out = ops(a)
    assert np.allclose(out.detach().numpy(), [0.87758255], atol=1e-6)  # test_operations.py:49-62
    out.sum().backward()
    assert np.allclose(a.grad.numpy(), [-np.sin(0.5)], atol=1e-6)

    n = 4
    circ = hea(n, 1, encoder=True)
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
    layer = MQLayer(sim.get_expectation_with_grad(mq.Hamiltonian(mq.QubitOperator("Z0"), mq.complex64), circ),
                    weight=np.linspace(0, 1, len(circ.ansatz_params_name)))
    x = torch.randn(5, n, requires_grad=True)
    y = layer(x)
    assert y.shape == (5, 1)
    y.sum().backward()
    assert x.grad.shape == (5, n) and layer.weight.grad.shape == (len(circ.ansatz_params_name),)
