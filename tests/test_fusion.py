"""Dense gate fusion on the host (mindquantum_amd.fusion): the fused blocks multiply to the circuit's
unitary, computed by the dense oracle from the gates' closed-form matrices (oracle/dense.py embeds
each gate into the full space independently of the fuser's block embedding); every gate lands in
exactly one block; block widths respect k; the C-ABI arrays carry the blocks in order."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import fusion
from mindquantum_amd.algorithm import layered_random_circuit
from mindquantum_amd.core import gates as G
from oracle import dense
from tests.test_gate_families import family_circuit


def _product(fc, n):
    v = np.eye(1 << n, dtype=np.complex128)
    for b in fc.blocks:
        v = dense.embed(b.matrix, list(b.qubits), [], n) @ v
    return v


@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
@pytest.mark.parametrize("seed", [0, 1])
def test_blocks_multiply_to_the_circuit(k, seed):
    n = 6 + seed
    circ = layered_random_circuit(n, 6, seed)
    fc = fusion.fuse(circ, k)
    assert fc.n_gates == len(circ)
    assert all(len(b.qubits) <= max(k, 2) for b in fc.blocks)
    assert np.max(np.abs(_product(fc, n) - dense.unitary(circ, {}, n))) < 1e-13


@pytest.mark.parametrize("k", [2, 3, 5])
def test_controls_multi_qubit_gates_and_parameters(k):
    """Controlled gates, 3-qubit Pauli rotations and shared parameters, resolved by a dict."""
    n = 5
    circ = family_circuit(n, 30, 3)
    names = list(circ.params_name)
    pr = dict(zip(names, np.random.default_rng(1).uniform(0, 2 * np.pi, size=len(names))))
    prims = mq.Circuit([p for g in circ for p in g.lowered()])
    fc = fusion.fuse(prims, k, pr)
    assert fc.n_gates == len(prims)
    assert np.max(np.abs(_product(fc, n) - dense.unitary(circ, pr, n))) < 1e-12


def test_parameterised_gate_needs_values_and_measurement_is_rejected():
    with pytest.raises(ValueError):
        fusion.fuse(mq.Circuit([G.RX("a").on(0)]), 2)
    with pytest.raises(ValueError):
        fusion.fuse(mq.Circuit([G.H.on(0), mq.Measure().on(0)]), 2)
    with pytest.raises(ValueError):
        fusion.fuse(mq.Circuit([G.H.on(0)]), 6)


def test_arrays_and_as_circuit():
    n = 6
    circ = layered_random_circuit(n, 4, 5)
    fc = fusion.fuse(circ, 3)
    nq, qs, mats = fc.arrays()
    assert nq.tolist() == [len(b.qubits) for b in fc.blocks]
    assert qs.tolist() == [q for b in fc.blocks for q in b.qubits]
    assert mats.size == sum(2 * 4 ** k for k in nq)
    first = fc.blocks[0].matrix
    assert np.array_equal(mats[: 2 * first.size].view(np.complex128).reshape(first.shape), first)
    tile = fc.as_circuit()
    assert np.max(np.abs(dense.unitary(tile, {}, n) - dense.unitary(circ, {}, n))) < 1e-13
    with pytest.raises(ValueError):
        fusion.fuse(layered_random_circuit(7, 5, 11), 5).as_circuit()
