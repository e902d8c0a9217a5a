"""Gate families lowered into the kernels' gate set (U3, FSim, Givens, SWAPalpha, Rxy, Rxz, Ryz,
RotPauliString, GroupedPauli, Power, constant Rn, Barrier).

Pins, independent of the lowering itself: every lowering's product against the gate's
closed-form matrix (the reference's matrix definitions, basicgate.py); the reference's own known
answer for U3 + FSim (test_simulator.py:520-562) through the CPU oracle; gradients of random
circuits over these families against the dense product-rule gradient, which differentiates the
closed-form matrices (diff_matrices) and never sees the lowering.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.core import gates as G
from oracle import cpu, dense

N = 4


def composite_cases():
    return [G.Rxy(0.7).on([1, 3], [0]), G.Rxz(-1.3).on([2, 0]), G.Ryz(2.1).on([0, 2], [3]),
            G.Givens(0.9).on([3, 1], [2]), G.SWAPalpha(0.37).on([0, 2], [1]), G.U3(0.3, 1.1, -0.7).on(2, [0, 1]),
            G.FSim(0.8, -2.2).on([1, 3], [0]), G.RotPauliString("XYZ", 1.0).on([0, 1, 3]),
            G.RotPauliString("YIXZ", 0.4).on([3, 2, 1, 0]), G.RotPauliString("Y", 0.4).on([2], [1]),
            G.Rn(0.3, 0.4, -0.5).on(1, [2]), G.GroupedPauli("XZY").on([0, 1, 2]), G.Power(G.H, 0.5).on(1, [0])]


@pytest.mark.parametrize("idx", range(len(composite_cases())))
def test_lowering_equals_closed_form(idx):
    g = composite_cases()[idx]
    full = dense.embed(dense.gate_matrix(g, {}), g.obj_qubits, g.ctrl_qubits, N)
    lowered = dense.unitary(mq.Circuit(g.lowered()), {}, N)
    assert np.max(np.abs(full - lowered)) < 1e-13, g.name


def test_hermitian_of_multi_parameter_gates():
    for g in (G.U3(0.3, 1.1, -0.7).on(0), G.FSim(0.8, -2.2).on([0, 1]), G.Rn(0.3, 0.4, -0.5).on(0),
              G.Givens(0.5).on([0, 1]), G.SWAPalpha(0.2).on([1, 0]), G.Rxy(0.6).on([0, 1])):
        m = g.matrix({}) if g.param_resolvers() else g.matrix()
        h = g.hermitian()
        mh = h.matrix({}) if h.param_resolvers() else h.matrix()
        assert np.allclose(mh, m.conj().T, atol=1e-14), g.name


def _golden_circuit():
    circ = mq.Circuit() + G.U3("a", "b", 1.0).on(0) + G.U3("c", "d", 2.0).on(1) + G.X.on(0, 1)
    circ += G.FSim("e", -3.0).on([0, 1])
    return circ


def test_golden_multi_params_gate(golden):
    gv = golden["multi_params_gate"]
    circ = _golden_circuit()
    names = list(circ.params_name)
    assert names == ["a", "b", "c", "d", "e"]
    p0 = np.array(gv["params"])
    qs_exp = np.array(gv["qs_re"]) + 1j * np.array(gv["qs_im"])
    for dt in (np.complex64, np.complex128):
        qs = cpu.apply_circuit(cpu.zero_state(2, dt), ir.lower(circ, names), p0)
        assert np.allclose(qs, qs_exp, atol=gv["atol"])
        f, g = cpu.grad_batch(2, dt, ir.lower(circ, names), [mq.Hamiltonian(mq.QubitOperator("X0 Y1")).terms_ir()],
                              np.zeros((1, 0)), p0)
        assert np.allclose(f, gv["f"], atol=gv["atol"])
        assert np.allclose(g[0, 0], gv["g"], atol=gv["atol"])
    # the dense simulation, from the closed-form matrices, agrees too
    f_d, g_d = dense.gradient(circ, dict(zip(names, p0)), names, 2, mq.QubitOperator("X0 Y1"))
    assert np.allclose(dense.state(circ, dict(zip(names, p0)), 2), qs_exp, atol=gv["atol"])
    assert np.allclose(f_d, gv["f"], atol=gv["atol"]) and np.allclose(g_d, gv["g"], atol=gv["atol"])


def family_circuit(n, depth, seed):
    """Random circuit over the lowered families with shared, scaled and offset parameters."""
    rng = np.random.default_rng(seed)
    circ = mq.Circuit()
    k = 0

    def pr():
        nonlocal k
        k += 1
        d = {f"t{k}": float(rng.uniform(0.5, 1.5))}
        if k > 1 and rng.random() < 0.3:
            d[f"t{int(rng.integers(1, k))}"] = float(rng.uniform(-1, 1))
        return mq.ParameterResolver(d, const=float(rng.uniform(-0.3, 0.3)))

    for _ in range(depth):
        qs = [int(q) for q in rng.permutation(n)]
        ctrl = [qs[3]] if n > 3 and rng.random() < 0.3 else []
        kind = int(rng.integers(9))
        if kind == 0:
            circ += G.U3(pr(), pr(), pr()).on(qs[0], ctrl)
        elif kind == 1:
            circ += G.FSim(pr(), pr()).on(qs[:2], ctrl)
        elif kind == 2:
            circ += G.Givens(pr()).on(qs[:2], ctrl)
        elif kind == 3:
            circ += G.SWAPalpha(pr()).on(qs[:2], ctrl)
        elif kind == 4:
            circ += (G.Rxy, G.Rxz, G.Ryz)[int(rng.integers(3))](pr()).on(qs[:2], ctrl)
        elif kind == 5:
            s = "".join("XYZ"[int(rng.integers(3))] for _ in range(3))
            circ += G.RotPauliString(s, pr()).on(qs[:3], ctrl)
        elif kind == 6:
            circ += G.RY(pr()).on(qs[0], ctrl)
        elif kind == 7:
            circ += G.H.on(qs[0])
        else:
            circ += G.X.on(qs[0], qs[1])
    return circ


@pytest.mark.parametrize("seed", range(4))
def test_family_gradients_match_dense(seed):
    n = 4 + seed % 2
    circ = family_circuit(n, 24, seed)
    op = mq.QubitOperator("Z0 X1") + 0.7 * mq.QubitOperator(f"Y{n - 1}") - 0.4 * mq.QubitOperator("X0 Z2 Y3")
    names = list(circ.params_name)
    vals = np.random.default_rng(seed).uniform(0, 2 * np.pi, size=len(names))
    f_ref, g_ref = dense.gradient(circ, dict(zip(names, vals)), names, n, op)
    f, g = cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()],
                          np.zeros((1, 0)), vals)
    assert np.allclose(f[0, 0], f_ref, atol=1e-10)
    assert np.allclose(g[0, 0].real, g_ref, atol=1e-9)


def test_parameterized_rn_lowers_to_a_parameterised_matrix():
    """Rn with a variable parameter runs as a parameterised matrix (per-sample values from its closed
    form); a constant Rn stays a fixed matrix."""
    circ = mq.Circuit([G.Rn("a", 0.1, 0.2).on(0), G.Rn(0.3, 0.1, 0.2).on(1)])
    kinds = list(ir.lower(circ, ["a"])[0]["kind"])
    assert kinds == [_native.GATE_KIND["PMATRIX"], _native.GATE_KIND["PMATRIX_TAP"], _native.GATE_KIND["PMATRIX_TAP"],
                     _native.GATE_KIND["MATRIX"]]
