"""Parameterised custom gates (gene_univ_parameterized_gate, gene_univ_two_params_gate) and
parameterised Rn on the host: the reference's golden state of generate_test_circuit
(test_simulator.py:227-293), the closed forms of the new derivative matrices against central finite
differences, and the lowering into MQA_GATE_PMATRIX / _TAP records with their per-sample table
(CPU only; the GPU runs them in test_gpu_custom_gates.py).
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.algorithm import qft
from mindquantum_amd.core import gates as G
from oracle import dense


def rx_matrix_generator(x):
    return np.array([[np.cos(x / 2), -1j * np.sin(x / 2)], [-1j * np.sin(x / 2), np.cos(x / 2)]])


def rx_diff_matrix_generator(x):
    return np.array([[np.sin(x / 2), 1j * np.cos(x / 2)], [1j * np.cos(x / 2), np.sin(x / 2)]]) / -2


def generate_test_circuit():
    """The reference's all-gate circuit (test_simulator.py:227-261) in this package's API."""
    circuit = mq.Circuit()
    circuit += mq.UN(G.H, 3)
    circuit.x(0).y(1).z(2)
    circuit += G.SWAP.on([0, 2], 1)
    circuit += mq.UN(G.X, 3)
    circuit += G.ISWAP.on([0, 1], 2)
    circuit.rx(1.2, 0).ry(2.3, 1).rz(3.4, 2)
    circuit.x(0, 1).x(1, 2).x(0, 2)
    circuit += G.PhaseShift(1.3).on(0, [1, 2])
    circuit += mq.UN(G.H, 3)
    circuit += mq.UN(G.S, 3)
    circuit += qft(range(3))
    circuit += mq.gene_univ_parameterized_gate("fake_x", rx_matrix_generator, rx_diff_matrix_generator)("a").on(0)
    circuit += G.RX("b").on(1, 2)
    circuit += G.RX("c").on(2, 0)
    circuit += mq.UN(G.H, 3)
    circuit += mq.UN(G.T, 3)
    circuit += G.UnivMathGate("fake_Rxx", G.Rxx(2.4).matrix()).on([0, 1])
    circuit += G.Ryy(4.6).on([1, 2])
    circuit += G.Rzz(6.8).on([0, 2])
    circuit += G.UnivMathGate("myX", G.X.matrix()).on(0)
    circuit += G.Power(G.X, 1.2).on(1)
    return circuit


def test_golden_all_gate_circuit_dense(golden):
    gv = golden["all_gate_circuit"]
    circ = generate_test_circuit()
    assert list(circ.params_name) == ["a", "b", "c"]
    qs = dense.state(circ, gv["params"], 3)
    assert np.allclose(qs, np.array(gv["qs_re"]) + 1j * np.array(gv["qs_im"]), atol=gv["atol"])


def _fd(fn, x, h=1e-6):
    return (fn(x + h) - fn(x - h)) / (2 * h)


def test_derivative_matrices_match_finite_differences():
    fx = mq.gene_univ_parameterized_gate("fx", rx_matrix_generator, rx_diff_matrix_generator)("a").on(0)
    for g in (fx, fx.hermitian()):
        assert np.allclose(g.diff_matrix({"a": 0.7}), _fd(lambda t: g.matrix({"a": t}), 0.7), atol=1e-8)
    rn = G.Rn("p", "q", "r").on(0)
    v = {"p": 0.3, "q": -1.1, "r": 0.8}
    for j, name in enumerate("pqr"):
        num = _fd(lambda t: rn.matrix({**v, name: t}), v[name])
        assert np.allclose(rn.diff_matrices(v)[j], num, atol=1e-8)
    zero = {"p": 0.0, "q": 0.0, "r": 0.0}
    for j, name in enumerate("pqr"):
        assert np.allclose(rn.diff_matrices(zero)[j], _fd(lambda t: rn.matrix({**zero, name: t}), 0.0), atol=1e-7)

    def m2(a, b):
        return np.array([[np.cos(a), -np.exp(1j * b) * np.sin(a)], [np.exp(-1j * b) * np.sin(a), np.cos(a)]])

    def d2a(a, b):
        return np.array([[-np.sin(a), -np.exp(1j * b) * np.cos(a)], [np.exp(-1j * b) * np.cos(a), -np.sin(a)]])

    def d2b(a, b):
        return np.array([[0, -1j * np.exp(1j * b) * np.sin(a)], [-1j * np.exp(-1j * b) * np.sin(a), 0]])

    two = mq.gene_univ_two_params_gate("two", m2, d2a, d2b)("u", "w").on(1)
    d = two.diff_matrices({"u": 0.4, "w": 1.7})
    assert np.allclose(d[0], _fd(lambda t: two.matrix({"u": t, "w": 1.7}), 0.4), atol=1e-8)
    assert np.allclose(d[1], _fd(lambda t: two.matrix({"u": 0.4, "w": t}), 1.7), atol=1e-8)


def test_lowering_records_and_table():
    fx = mq.gene_univ_parameterized_gate("fx", rx_matrix_generator, rx_diff_matrix_generator)
    circ = mq.Circuit() + fx({"a": 2.0, "b": 1.0}).on(0, 1) + G.Rn("p", "q", 0.3).on(1) + G.RX("b").on(0)
    names = ["a", "b", "p", "q"]
    layout = []
    gates = ir.lower(circ, names, layout)[0]
    assert list(gates["kind"]) == [40, 40, 41, 41, 32]
    assert [(off, k, m) for _, off, k, m in layout] == [(0, 1, 1), (24, 1, 3)]
    vals = np.array([[0.1, 0.2, 0.3, 0.4], [0.5, 0.6, 0.7, 0.8]])
    table, stride = ir.custom_table(layout, names, vals)
    assert stride == 64 and table.shape == (2, 64)
    for r in range(2):
        pr = dict(zip(names, vals[r]))
        g0 = fx({"a": 2.0, "b": 1.0}).on(0).matrix(pr)
        blk = table[r, 0:24:2] + 1j * table[r, 1:24:2]
        assert np.allclose(blk[:4].reshape(2, 2), g0) and np.allclose(blk[4:8].reshape(2, 2), g0.conj().T)
        m0 = fx({"a": 2.0, "b": 1.0}).on(0).diff_matrix(pr) @ g0.conj().T
        assert np.allclose(blk[8:].reshape(2, 2), m0)


def test_wide_custom_gate_is_rejected_at_lowering():
    def m3(x):
        return np.eye(8, dtype=complex) * np.exp(1j * x)

    g3 = mq.gene_univ_parameterized_gate("g3", m3, lambda x: 1j * m3(x))
    with pytest.raises(ValueError):
        ir.lower(mq.Circuit() + g3("a").on([0, 1, 2]), ["a"])


def test_custom_matrix_table_abi_checks():
    fx = mq.gene_univ_parameterized_gate("fx", rx_matrix_generator, rx_diff_matrix_generator)
    nc = ir.NativeCircuit(mq.Circuit() + fx("a").on(0), ["a"])
    short = np.zeros(10)
    assert _native.lib().mqa_circuit_set_custom_matrices(nc.handle, 1, 10, _native.ptr(short)) == _native.MQA_ERR_ARG
    nc.set_custom(np.array([[0.5]]))
