"""Host-side logic: parameter resolver, circuit bookkeeping, operators, IR lowering."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.core import gates as G
from mindquantum_amd.simulator.utils import _thread_balance


def test_parameter_resolver_algebra_and_flags():
    pr = mq.ParameterResolver({"a": 0.3, "b": 0.5}, const=0.1)
    assert pr.combination({"a": 2, "b": 4}) == pytest.approx(0.1 + 0.6 + 2.0)
    pr.no_grad_part("a")
    assert pr.no_grad_parameters == ["a"] and pr.requires_grad_parameters == ["b"]
    pr2 = -pr * 2
    assert pr2["a"] == pytest.approx(-0.6) and pr2.const == pytest.approx(-0.2)
    assert mq.ParameterResolver("x").as_encoder().encoder_parameters == ["x"]
    assert mq.ParameterResolver(3).is_const()


def test_circuit_names_and_hermitian():
    enc = mq.Circuit().rx("x0", 0).rx("x1", 1).as_encoder()
    circ = enc + mq.Circuit().ry("a", 0).rz({"b": 2, "a": 1}, 1).x(1, 0)
    assert circ.params_name == ["x0", "x1", "a", "b"]
    assert circ.encoder_params_name == ["x0", "x1"]
    assert circ.ansatz_params_name == ["a", "b"]
    assert circ.n_qubits == 2
    herm = circ.hermitian()
    assert [g.name for g in herm] == ["X", "RZ", "RY", "RX", "RX"]
    assert herm[1].coeff["b"] == -2
    s = G.S.on(0)
    assert s.hermitian().ir_kind == "Sdag" and s.hermitian().hermitian().ir_kind == "S"


def test_gate_matrices_are_unitary_and_diff_is_derivative():
    for g in [G.RX(0.3), G.RY(0.3), G.RZ(0.3), G.Rxx(0.3), G.Ryy(0.3), G.Rzz(0.3), G.PhaseShift(0.3),
              G.GlobalPhase(0.3)]:
        m = g.matrix()
        assert np.allclose(m.conj().T @ m, np.eye(m.shape[0]))
        h = 1e-6
        fd = (type(g)(0.3 + h).matrix() - type(g)(0.3 - h).matrix()) / (2 * h)
        assert np.allclose(g.diff_matrix(), fd, atol=1e-8)
    # reference conventions (basicgate.py)
    assert np.allclose(G.RY(np.pi).matrix(), [[0, -1], [1, 0]], atol=1e-12)
    assert np.allclose(G.Ryy(np.pi).matrix()[0, 3], 1j)
    assert np.allclose(G.GlobalPhase(0.5).matrix(), np.exp(-0.5j) * np.eye(2))


def test_qubit_operator_and_hamiltonian_ir():
    op = mq.QubitOperator("X0 Y1", 0.5) + mq.QubitOperator("Z2") * 2 - 1.0
    terms = mq.Hamiltonian(op).terms_ir()
    assert len(terms) == 3
    t = {(int(r["x_mask"]), int(r["y_mask"]), int(r["z_mask"])): float(r["coeff"]) for r in terms}
    assert t[(1, 2, 0)] == 0.5 and t[(0, 0, 4)] == 2.0 and t[(0, 0, 0)] == -1.0
    prod = mq.QubitOperator("X0") * mq.QubitOperator("Y0")
    assert prod.terms == {((0, "Z"),): 1j}
    with pytest.raises(ValueError):
        mq.Hamiltonian(mq.QubitOperator("X0", 1j))


def test_ir_struct_layout():
    circ = mq.Circuit().rx({"a": 2.0}, 3, [0, 5]).swap([1, 2]) + G.UnivMathGate("u", np.eye(4)).on([0, 1])
    gates, tp, tc, mats, n_mats, rg = ir.lower(circ, ["a"])
    assert gates.dtype.itemsize == 64
    assert gates[0]["kind"] == _native.GATE_KIND["RX"] and gates[0]["obj"][0] == 3
    assert int(gates[0]["ctrl_mask"]) == (1 << 0) | (1 << 5)
    assert list(tp) == [0] and list(tc) == [2.0]
    assert gates[1]["kind"] == _native.GATE_KIND["SWAP"] and gates[1]["n_obj"] == 2
    assert gates[2]["kind"] == _native.GATE_KIND["MATRIX"] and n_mats == 16
    assert list(rg) == [1]


def test_thread_balance_matches_reference_rule():
    assert _thread_balance(1024, 1, None) == (1024, 1)
    assert _thread_balance(10, 2, 4) == (4, 1)
    assert _thread_balance(2, 4, 3) == (1, 3)
