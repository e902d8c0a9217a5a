"""Pins of the CPU oracle (oracle/mqvector_oracle.cpp) before it is trusted as the GPU checker.

Each check is independent of the oracle's own formulas: reference golden values, closed
forms, the dense numpy simulation (oracle/dense.py, built from Kronecker products), central
finite differences, the parameter-shift rule and the hermitian round trip.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import ir
from mindquantum_amd.core import gates as G
from oracle import cpu, dense
from tests.conftest import hea, params_for, random_circuit, random_hamiltonian


def _ham_terms(op):
    return mq.Hamiltonian(op).terms_ir()


def oracle_state(circ, n, values, dtype=np.complex128, init=None):
    names = list(circ.params_name)
    arrays = ir.lower(circ, names)
    qs = cpu.zero_state(n, dtype) if init is None else np.array(init, dtype=dtype)
    vec = np.array([values[k] for k in names], dtype=np.float64)
    return cpu.apply_circuit(qs, arrays, vec)


def oracle_grad(circ, n, op, enc, ans, dtype=np.complex128, batch_threads=1):
    enc_names = list(circ.encoder_params_name)
    ans_names = list(circ.ansatz_params_name)
    arrays = ir.lower(circ, enc_names + ans_names)
    return cpu.grad_batch(n, dtype, arrays, [_ham_terms(op)], enc, ans, batch_threads)


# ---------------------------------------------------------------- golden values
def test_golden_ry_docstring(golden):
    gv = golden["ry_z0_docstring"]
    circ = mq.Circuit().ry("a", 0)
    for dt in (np.complex64, np.complex128):
        f, g = oracle_grad(circ, 1, mq.QubitOperator("Z0"), np.zeros((1, 0)), np.array([gv["a"]]), dt)
        assert np.allclose(f, gv["f"], atol=1e-6)
        assert np.allclose(g, gv["g"], atol=1e-6)


def test_golden_rx_ops(golden):
    gv = golden["rx_z0_ops"]
    circ = mq.Circuit().rx("a", 0)
    f, _ = oracle_grad(circ, 1, mq.QubitOperator("Z0"), np.zeros((1, 0)), np.array([gv["a"]]))
    assert np.allclose(f.real, gv["f"], atol=1e-7)


def _mul_qubit_circuit(x):
    diag = np.diag([np.exp(1j * 2 * (k + 1) * x) for k in range(8)])
    circ = mq.UN(G.H, 3)
    circ += G.UnivMathGate("q", diag).on([0, 1, 2])
    return circ


def test_golden_mul_qubit_gate(golden):
    gv = golden["mul_qubit_gate"]
    terms = _ham_terms(mq.QubitOperator("X0"))

    def f_of(x):
        qs = oracle_state(_mul_qubit_circuit(x), 3, {})
        return cpu.expectation(qs, qs, terms)

    f = f_of(gv["x"])
    assert np.allclose(f, gv["f"], atol=1e-7)
    h = gv["fd_step"]
    g = (f_of(gv["x"] + h) - f_of(gv["x"] - h)).real / (2 * h)
    assert np.allclose(g, gv["g"], atol=1e-5)


# ---------------------------------------------------------------- dense simulation
@pytest.mark.parametrize("seed", range(6))
def test_states_match_dense(seed):
    n = 3 + seed % 4
    circ = random_circuit(n, 40, seed)
    names, vals = params_for(circ, seed)
    values = dict(zip(names, vals))
    ref = dense.state(circ, values, n)
    got = oracle_state(circ, n, values)
    assert 1 - abs(np.vdot(ref, got)) < 1e-12
    got64 = oracle_state(circ, n, values, np.complex64)
    assert np.max(np.abs(got64 - ref)) < 2e-5


@pytest.mark.parametrize("seed", range(4))
def test_hamiltonian_and_expectation_match_dense(seed):
    n = 4 + seed
    op = random_hamiltonian(n, 12, seed)
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    phi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    hm = dense.hamiltonian_matrix(op, n)
    terms = _ham_terms(op)
    assert np.allclose(cpu.apply_hamiltonian(psi, terms), hm @ psi, atol=1e-10)
    assert np.allclose(cpu.expectation(phi, psi, terms), np.vdot(phi, hm @ psi), atol=1e-9)
    assert np.allclose(cpu.vdot(phi, psi), np.vdot(phi, psi), atol=1e-10)


@pytest.mark.parametrize("seed", range(5))
def test_gradient_matches_dense(seed):
    n = 3 + seed % 3
    circ = random_circuit(n, 30, 100 + seed, ctrl_prob=0.3)
    op = random_hamiltonian(n, 8, seed)
    names, vals = params_for(circ, seed)
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    f, g = oracle_grad(circ, n, op, np.zeros((1, 0)), vals)
    assert np.allclose(f[0, 0], f_ref, atol=1e-10)
    assert np.allclose(g[0, 0].real, g_ref, atol=1e-9)
    assert np.allclose(g[0, 0].imag, 0)


def test_gradient_central_fd_and_encoder_batch():
    n = 5
    circ = hea(n, 2, encoder=True)
    op = mq.QubitOperator("Z0 Z1") + 0.5 * mq.QubitOperator("X2") - 0.3 * mq.QubitOperator("Y3 Y4")
    rng = np.random.default_rng(3)
    enc = rng.normal(size=(3, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    f, g = oracle_grad(circ, n, op, enc, ans, batch_threads=2)
    terms = _ham_terms(op)
    enc_names = list(circ.encoder_params_name)
    ans_names = list(circ.ansatz_params_name)
    h = 1e-5
    for b in range(3):
        base = dict(zip(enc_names, enc[b]))
        base.update(zip(ans_names, ans))
        for k, name in enumerate(enc_names + ans_names):
            if k % 3:
                continue
            vp, vm = dict(base), dict(base)
            vp[name] += h
            vm[name] -= h
            qp, qm = oracle_state(circ, n, vp), oracle_state(circ, n, vm)
            fd = (cpu.expectation(qp, qp, terms) - cpu.expectation(qm, qm, terms)).real / (2 * h)
            assert abs(fd - g[b, 0, k].real) < 1e-6


@pytest.mark.parametrize("kind", ["RX", "RY", "RZ", "Rzz", "Rxx", "Ryy"])
def test_parameter_shift_rule(kind):
    """d/dth <H> = (f(th + pi/2) - f(th - pi/2)) / 2 for exp(-i th/2 P) gates (vector_state.tpp:1157-1171)."""
    n = 3
    circ = mq.Circuit().h(0).h(1).ry(0.3, 2).x(2, 0)
    cls = {"RX": G.RX, "RY": G.RY, "RZ": G.RZ, "Rzz": G.Rzz, "Rxx": G.Rxx, "Ryy": G.Ryy}[kind]
    obj = [1] if kind in ("RX", "RY", "RZ") else [1, 2]
    circ += cls("t").on(obj)
    circ.rx(0.7, 0)
    op = mq.QubitOperator("Z1") + mq.QubitOperator("X0 Y2", 0.4)
    terms = _ham_terms(op)
    th = 0.81

    def f_of(v):
        q = oracle_state(circ, n, {"t": v})
        return cpu.expectation(q, q, terms).real

    shift = (f_of(th + np.pi / 2) - f_of(th - np.pi / 2)) / 2
    _, g = oracle_grad(circ, n, op, np.zeros((1, 0)), np.array([th]))
    assert abs(shift - g[0, 0, 0].real) < 1e-10


def test_hermitian_round_trip():
    """apply circ then circ.hermitian() returns the initial state (test_simulator.py:125-167)."""
    n = 5
    circ = random_circuit(n, 60, 7)
    names, vals = params_for(circ, 7)
    values = dict(zip(names, vals))
    rng = np.random.default_rng(1)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    psi /= np.linalg.norm(psi)
    out = oracle_state(circ + circ.hermitian(), n, values, init=psi)
    assert np.allclose(out, psi, atol=1e-12)


def test_c64_within_tolerance_of_c128():
    n = 8
    circ = hea(n, 3)
    op = random_hamiltonian(n, 10, 2)
    names, vals = params_for(circ, 2)
    f64, g64 = oracle_grad(circ, n, op, np.zeros((1, 0)), vals, np.complex64)
    f128, g128 = oracle_grad(circ, n, op, np.zeros((1, 0)), vals, np.complex128)
    assert abs(f64[0, 0] - f128[0, 0]) < 1e-5
    assert np.max(np.abs(g64 - g128)) < 1e-4


def test_planted_mistake_is_caught():
    """A dropped term in the Hamiltonian application must fail the dense comparison."""
    n = 4
    op = random_hamiltonian(n, 6, 11)
    terms = _ham_terms(op)
    rng = np.random.default_rng(0)
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    wrong = cpu.apply_hamiltonian(psi, terms[:-1])
    assert not np.allclose(wrong, dense.hamiltonian_matrix(op, n) @ psi, atol=1e-6)
