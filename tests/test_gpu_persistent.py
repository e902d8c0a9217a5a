"""Persistent compiled back-sweep passes (option ``jit_persistent``, complex64): a block walks
(tile, sample) units with a stride of its grid and loads the next unit's tiles while it computes
the current one.  Gradients must match the oracle (bars of test_gpu_parity.py), with the grid
capped so that every block runs many units (the prefetch path), with exact and phase-free plans,
narrow and default tiles, checkpoints, and controls that cross the tile.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.core import gates as G
from tests.conftest import hea, random_hamiltonian
from tests.test_gpu_parity import assert_grad, oracle_grad

pytestmark = pytest.mark.gpu

KEYS = ("jit", "jit_min_amps", "phase_free", "checkpoints", "jit_persistent", "jit_persistent_blocks")


@pytest.fixture
def options():
    saved = {k: _native.get_option(k) for k in KEYS}
    _native.set_option("jit", 2)
    _native.set_option("jit_min_amps", 0)
    _native.set_option("jit_persistent", 1)
    yield
    for k, v in saved.items():
        _native.set_option(k, v)


def circuit(n):
    circ = hea(n, 3, encoder=True)
    circ += G.RY("cy").on(1, n - 1)          # control outside most tiles
    circ += G.X.on(n - 2, [0, n - 1])         # Toffoli
    circ += G.RZ({"cz": 1.0, "a0": 0.5}).on(3)
    return circ


@pytest.mark.parametrize("phase_free", [False, True])
@pytest.mark.parametrize("tile,blocks", [(9, 7), (12, 5), (12, 0)])
def test_persistent_gradient_matches_oracle(options, phase_free, tile, blocks):
    n, batch = 14, 5
    _native.set_option("phase_free", int(phase_free))
    _native.set_option("jit_persistent_blocks", blocks)
    circ = circuit(n)
    code = mq.dtype.native_code(mq.complex64)
    circ._mqa_tiles = {(code, True): tile}
    ops = [random_hamiltonian(n, 9, 4), mq.QubitOperator("X0 Y5") + 0.3 * mq.QubitOperator("Z13")]
    rng = np.random.default_rng(tile + blocks)
    enc = rng.normal(size=(batch, len(circ.encoder_params_name)))
    ans = rng.uniform(-2 * np.pi, 2 * np.pi, size=len(circ.ansatz_params_name))
    before = _native.get_option("jit_persistent_launches")
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, mq.complex64) for o in ops], circ)(enc, ans)
    assert _native.get_option("jit_persistent_launches") > before
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans, mq.complex64)
    assert np.max(np.abs(f - f_ref)) < 2e-4
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, mq.complex64)


def test_persistent_with_checkpoints(options):
    n, batch = 15, 3
    _native.set_option("checkpoints", 2)
    _native.set_option("jit_persistent_blocks", 11)
    circ = circuit(n)
    code = mq.dtype.native_code(mq.complex64)
    circ._mqa_tiles = {(code, True): 10, (code, False): 10}
    op = random_hamiltonian(n, 10, 9)
    rng = np.random.default_rng(5)
    enc = rng.normal(size=(batch, len(circ.encoder_params_name)))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
    f, g_enc, g_ans = sim.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex64), circ)(enc, ans)
    assert _native.get_option("checkpoints_used") > 0
    f_ref, g_ref = oracle_grad(circ, n, [op], enc, ans, mq.complex64)
    assert np.max(np.abs(f - f_ref)) < 2e-4
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, mq.complex64)
