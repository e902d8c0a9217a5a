"""Single-buffer compiled back-sweep passes (option ``jit_single_buffer``, complex64): psi and
lambda stay in registers between register groups and share one LDS tile buffer.  Gradients must
match the complex128 oracle (bars of tests/bars.py) with exact and phase-free plans, narrow and
default tiles, checkpoints, controls that cross the tile, and a dense 3-qubit block (which keeps the
two-buffer form for its circuit).
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.core import gates as G
from tests.bars import assert_f, coeff_sum
from tests.conftest import hea, random_hamiltonian
from tests.test_gpu_parity import assert_grad, oracle_grad

pytestmark = pytest.mark.gpu

KEYS = ("jit", "jit_min_amps", "phase_free", "checkpoints", "jit_single_buffer")


@pytest.fixture
def options():
    saved = {k: _native.get_option(k) for k in KEYS}
    _native.set_option("jit", 2)
    _native.set_option("jit_min_amps", 0)
    _native.set_option("jit_single_buffer", 1)
    yield
    for k, v in saved.items():
        _native.set_option(k, v)


def circuit(n):
    circ = hea(n, 3, encoder=True)
    circ += G.RY("cy").on(1, n - 1)
    circ += G.X.on(n - 2, [0, n - 1])
    circ += G.RZ({"cz": 1.0, "a0": 0.5}).on(3)
    return circ


def run(circ, n, ops, batch, seed, tiles=None):
    code = mq.dtype.native_code(mq.complex64)
    if tiles:
        circ._mqa_tiles = {(code, True): tiles[1], (code, False): tiles[0]}
    rng = np.random.default_rng(seed)
    enc = rng.normal(size=(batch, len(circ.encoder_params_name)))
    ans = rng.uniform(-2 * np.pi, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex64)
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, mq.complex64) for o in ops], circ)(enc, ans)
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans)
    for h, o in enumerate(ops):
        assert_f(f[:, h], f_ref[:, h], mq.complex64, coeff_sum(o))
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, mq.complex64)


@pytest.mark.parametrize("phase_free", [False, True])
@pytest.mark.parametrize("tile", [8, 11, 12, 13])
def test_single_buffer_gradient_matches_oracle(options, phase_free, tile):
    """Tiles of one partial wave and of 2, 4 and 8 waves."""
    n = 14
    _native.set_option("phase_free", int(phase_free))
    before = _native.get_option("jit_single_buffer_launches")
    run(circuit(n), n, [random_hamiltonian(n, 9, 4), mq.QubitOperator("X0 Y5") + 0.3 * mq.QubitOperator("Z13")],
        4, tile, tiles=(12, tile))
    assert _native.get_option("jit_single_buffer_launches") > before


def test_single_buffer_with_checkpoints(options):
    n = 15
    _native.set_option("checkpoints", 2)
    before = _native.get_option("jit_single_buffer_launches")
    run(circuit(n), n, [random_hamiltonian(n, 10, 9)], 3, 5, tiles=(10, 10))
    assert _native.get_option("checkpoints_used") > 0
    assert _native.get_option("jit_single_buffer_launches") > before


def test_dense_block_keeps_two_buffers(options):
    n = 13
    rng = np.random.default_rng(8)
    m = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
    circ = circuit(n) + G.UnivMathGate("u3", m).on([2, 5, 9])
    before = _native.get_option("jit_single_buffer_launches")
    run(circ, n, [random_hamiltonian(n, 8, 2)], 3, 9)
    assert _native.get_option("jit_single_buffer_launches") == before
