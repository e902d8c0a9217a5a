"""Dense blocks on the GPU (mqa_sim_apply_dense_blocks and the tile passes) against the complex128
oracle: one random dense k-qubit unitary (k = 1..5) on scattered qubits against the dense oracle's
embedding, the complex64 5-qubit block on the FP32 matrix cores and on the vector ALUs, fused
circuits (k = 2..5) in both forms against the oracle's state of the unfused circuit, and the host's
argument checks (nothing launched)."""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, fusion, ir
from mindquantum_amd.algorithm import layered_random_circuit
from oracle import cpu, dense
from tests.bars import assert_state

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def _random_unitary(k, rng):
    a = rng.normal(size=(1 << k, 1 << k)) + 1j * rng.normal(size=(1 << k, 1 << k))
    q, r = np.linalg.qr(a)
    return q * (np.diag(r) / np.abs(np.diag(r)))


def _random_state(n, rng):
    psi = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return psi / np.linalg.norm(psi)


@pytest.fixture
def mfma_option():
    saved = _native.get_option("dense_mfma")
    yield
    _native.set_option("dense_mfma", saved)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("k", [1, 2, 3, 4, 5])
def test_dense_block_matches_embedding(dtype, k, mfma_option):
    rng = np.random.default_rng(10 + k)
    n = 11
    for mfma in (1, 0) if k == 5 else (1,):
        _native.set_option("dense_mfma", mfma)
        qubits = [int(q) for q in rng.permutation(n)[:k]]  # unsorted: bit b of the matrix <-> qubits[b]
        u = _random_unitary(k, rng)
        psi = _random_state(n, rng)
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        sim.set_qs(psi)
        fc = fusion.FusedCircuit([fusion.DenseBlock(qubits, u, [])], k)
        fc.apply(sim)
        assert_state(sim.get_qs(), dense.embed(u, qubits, [], n) @ psi, dtype)


@pytest.mark.parametrize("n", [10, 14])
def test_mfma_and_valu_forms_agree(n, mfma_option):
    """The complex64 5-qubit sweep on the matrix cores and on the VALU, with the low qubits among the
    targets (qubit 0 is row bit 0) and without, both against the complex128 embedding."""
    rng = np.random.default_rng(n)
    for qubits in ([0, 3, 1, n - 1, 6], [n - 1, n - 4, 5, 7, n - 2]):
        u = _random_unitary(5, rng)
        psi = _random_state(n, rng)
        ref = dense.embed(u, qubits, [], n) @ psi if n <= 12 else None
        out = []
        for mfma in (1, 0):
            _native.set_option("dense_mfma", mfma)
            sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
            sim.set_qs(psi)
            _native.profile_reset()
            _native.set_option("profile", 1)
            fusion.FusedCircuit([fusion.DenseBlock(qubits, u, [])], 5).apply(sim)
            _native.set_option("profile", 0)
            assert _native.profile("dense_sweep_mfma" if mfma else "dense_sweep")[0] == 1
            out.append(sim.get_qs())
        if ref is None:  # apply by tensor contraction in numpy for the 14-qubit case
            t = psi.reshape([2] * n)
            axes = [n - 1 - q for q in reversed(qubits)]  # numpy axis of qubit q is n - 1 - q
            ut = u.reshape([2] * 10)
            t = np.tensordot(ut, t, axes=(list(range(5, 10)), axes))
            t = np.moveaxis(t, list(range(5)), axes)
            ref = t.reshape(-1)
        for o in out:
            assert_state(o, ref, mq.complex64)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("k", [2, 3, 4, 5])
def test_fused_circuit_state(dtype, k):
    n = 12
    circ = layered_random_circuit(n, 12, 40 + k)
    ref = cpu.apply_circuit(cpu.zero_state(n, np.complex128), ir.lower(circ, []), np.zeros(0))
    fc = fusion.fuse(circ, k)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    fc.apply(sim)
    assert_state(sim.get_qs(), ref, dtype)
    if k <= 4:
        sim2 = mq.Simulator("mqvector_amd", n, dtype=dtype)
        sim2.apply_circuit(fc.as_circuit())
        assert_state(sim2.get_qs(), ref, dtype)


def test_bad_blocks_raise_before_launch():
    sim = mq.Simulator("mqvector_amd", 6, dtype=mq.complex64)
    psi = _random_state(6, np.random.default_rng(0))
    sim.set_qs(psi)
    u = _random_unitary(2, np.random.default_rng(1))
    lib = _native.lib()
    good = fusion.FusedCircuit([fusion.DenseBlock([0, 1], u, [])], 2)
    for nq, qs in (([2], [0, 0]), ([2], [0, 6]), ([6], [0, 1, 2, 3, 4, 5]), ([0], [])):
        nq = np.array(nq, dtype=np.int32)
        qs = np.array(qs + [0] * 6, dtype=np.int32)
        mats = np.zeros(2 * 4 ** 6)
        with pytest.raises(ValueError):
            _native.check(lib.mqa_sim_apply_dense_blocks(sim.backend._h, 1, _native.ptr(nq), _native.ptr(qs),
                                                         _native.ptr(mats)))
    # nothing ran: the state is unchanged, and a good block still applies
    np.testing.assert_array_equal(sim.get_qs(), psi.astype(np.complex64))
    good.apply(sim)
    assert_state(sim.get_qs(), dense.embed(u, [0, 1], [], 6) @ psi, mq.complex64)
