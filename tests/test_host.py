"""Host-side logic: parameter resolver, circuit bookkeeping, operators, IR lowering."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.core import gates as G
from mindquantum_amd.simulator.utils import check_parallel_worker


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


def test_circuit_name_views_follow_every_change():
    """The name views are kept between reads; any change to a circuit, gate or resolver renews them."""
    circ = mq.Circuit().rx("a", 0).ry("b", 1)
    first = circ.params_name
    assert first == ["a", "b"] and circ.ansatz_params_name == ["a", "b"] and circ.encoder_params_name == []
    first.append("z")  # a caller's copy: the kept view is not touched
    assert circ.params_name == ["a", "b"]
    circ += mq.Circuit().rz("c", 0)
    assert circ.params_name == ["a", "b", "c"]
    circ[0] = G.RX("d").on(0)
    assert circ.params_name == ["d", "b", "c"]
    circ.as_encoder()
    assert circ.encoder_params_name == ["d", "b", "c"] and circ.ansatz_params_name == []
    other = mq.Circuit().ry("e", 0)
    assert other.params_name == ["e"] and circ.params_name == ["d", "b", "c"]
    other[0].coeff.as_encoder()  # a resolver's flag on another circuit's gate
    assert other.encoder_params_name == ["e"]


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


def test_parallel_worker_validation():
    check_parallel_worker(None)
    check_parallel_worker(4)
    with pytest.raises(TypeError):
        check_parallel_worker(2.0)
    with pytest.raises(ValueError):
        check_parallel_worker(0)


def test_native_cache_sees_circuit_and_parameter_changes():
    """A native circuit is reused only while the circuit's lowered content is unchanged: making a
    parameter no-grad, replacing a gate, or inserting then removing one builds a new native circuit
    (or reuses the old one when the content is back to what it was)."""
    from mindquantum_amd.ir import NativeCache

    cache = NativeCache(8)
    c = mq.Circuit().rx("a", 0).ry("b", 1)
    names = ["a", "b"]
    n0 = cache.get(c, names)
    assert cache.get(c, names) is n0
    kinds0 = list(n0.arrays[0]["kind"])
    c.no_grad()
    n1 = cache.get(c, names)
    assert n1 is not n0 and list(n1.arrays[5]) == [0, 0]
    c.requires_grad()
    c[0] = G.RZ("a").on(0)
    n2 = cache.get(c, names)
    assert list(n2.arrays[0]["kind"]) != kinds0 and list(n2.arrays[5]) == [1, 1]
    c.insert(1, G.H.on(1))
    n3 = cache.get(c, names)
    assert len(n3.arrays[0]) == 3
    c.pop(1)
    n4 = cache.get(c, names)
    assert n4.digest == n2.digest and len(n4.arrays[0]) == 2
    c[0].coeff.no_grad_part("a")
    assert list(cache.get(c, names).arrays[5]) == [0, 1]
    # entries die with their circuit; the size bound holds
    del c
    import gc

    gc.collect()
    assert len(cache) == 0
    keep = [mq.Circuit().rx(f"p{i}", 0) for i in range(12)]
    for k in keep:
        cache.get(k, list(k.params_name))
    assert len(cache) == 8


def test_grad_workspaces_never_outgrow_the_device():
    """Sequences of batched-gradient calls on one simulator, against a fake device of fixed
    capacity: each call's budget is 85% of (free + held) as mqa_sim_grad_batch computes it, the
    planned releases happen first and every allocation must then fit.  Before the plan released
    stale buffers, a one-Hamiltonian call that filled the budget followed by a two-Hamiltonian call
    (which keeps a copy of the forward state) allocated the copy on top of the oversized psi and
    lambda buffers."""
    cap = 256 << 30
    sb = 1 << 30
    side = 4096
    held = [0, 0, 0, 0]

    def call(n_hams, ck, batch):
        free = cap - sum(held)
        budget = int(0.85 * (free + sum(held)))
        chunk, keep, need, release = _native.grad_workspace_plan(budget, sb, side, n_hams, ck, batch, held)
        for i in range(4):
            if release[i]:
                held[i] = 0
        for i in range(4):
            if need[i] > held[i]:  # DevBuf::ensure frees the old buffer, then allocates
                held[i] = 0
                assert sum(held) + need[i] <= cap, f"allocation of {need[i]} does not fit"
                held[i] = need[i]
        return chunk, keep

    c1, k1 = call(1, 0, 4096)
    assert not k1 and c1 == int(0.85 * cap) // (2 * sb + side)
    c2, k2 = call(2, 0, 4096)
    assert k2 and c2 == int(0.85 * cap) // (3 * sb + side)
    c3, _ = call(2, 0, 4096)  # repeated: same plan, nothing released
    assert c3 == c2
    before = list(held)
    chunk, keep, need, release = _native.grad_workspace_plan(int(0.85 * cap), sb, side, 2, 0, 4096, held)
    assert not any(release) and need == before
    call(1, 4, 4096)  # checkpoints: the kept copy goes, the checkpoint buffer comes
    assert held[2] == 0 and held[3] > 0
    call(1, 0, 8)  # a small batch keeps the oversized psi / lambda buffers (they fit the budget)
    assert held[0] >= 8 * sb and held[3] == 0
    call(3, 0, 100000)
