"""GPU parity of the Hamiltonian engines (H|psi>, <psi|H|psi> and the adjoint gradient that starts
from them) against the CPU oracle and against each other.

The tile engine packs the flip masks of a Hamiltonian into passes over LDS-staged tiles; these
tests force multi-pass plans (small tiles, flips on high qubits), every weight kind (terms whose
sign depends on no tile bit, on thread bits only, on register bits), Y terms (imaginary
coefficients) and the fallback to the gather kernel when a flip mask fits no tile.

Bars (tests/bars.py) against the complex128 oracle on the same input state: 1e-10 (complex128) /
1e-5 (complex64) times max(1, sum |c|) for H psi and expectations; gradients 1e-10 / 1e-5 relative
to the largest |g_ref|.
"""
import contextlib

import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from oracle import cpu
from tests.bars import assert_g, f_tol
from tests.conftest import hea, random_hamiltonian

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


@contextlib.contextmanager
def options(**kv):
    try:
        for k, v in kv.items():
            _native.set_option(k, v)
        yield
    finally:
        for k in kv:
            _native.set_option(k, -1)


def random_state(n, seed, dtype):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    return (v / np.linalg.norm(v)).astype(dtype)


def tol_for(op, dtype):
    return f_tol(dtype, sum(abs(v) for v in op.terms.values()))


def ring_hamiltonian(n, seed):
    rng = np.random.default_rng(seed)
    op = mq.QubitOperator("", 0)
    for i in range(n):
        op += mq.QubitOperator(f"Z{i} Z{(i + 1) % n}", float(rng.uniform(-1, 1)))
        op += mq.QubitOperator(f"X{i}", float(rng.uniform(-1, 1)))
    return op


def check_apply(n, op, dtype, seed):
    ham = mq.Hamiltonian(op, dtype)
    qs = random_state(n, seed, dtype)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.set_qs(qs)
    sim.apply_hamiltonian(ham)
    got = sim.get_qs()
    ref = cpu.apply_hamiltonian(qs.astype(np.complex128), mq.Hamiltonian(op, mq.complex128).terms_ir())
    tol = tol_for(op, dtype)
    assert np.max(np.abs(got - ref)) < tol, f"n={n} max err {np.max(np.abs(got - ref))}"
    return got


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,seed,weight", [(6, 0, 3), (9, 1, 5), (15, 2, 4), (18, 3, 6)])
def test_apply_hamiltonian_random(n, seed, weight, dtype):
    check_apply(n, random_hamiltonian(n, 24, seed, max_weight=weight), dtype, seed)


@pytest.mark.parametrize("dtype", DTYPES)
def test_apply_hamiltonian_ring_multi_pass(dtype):
    """The 20-qubit ZZ-ring + X Hamiltonian: single-qubit flips on all 20 qubits need two passes."""
    check_apply(20, ring_hamiltonian(20, 7), dtype, 7)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("k", [6, 7, 9])
def test_small_tiles_force_many_passes(k, dtype):
    """Small tiles: many passes, and yz bits on thread and register bits in every pass."""
    n = 14
    op = random_hamiltonian(n, 40, 30 + k, max_weight=4)
    key = "ham_tile_qubits_c64" if dtype == mq.complex64 else "ham_tile_qubits_c128"
    with options(**{key: k}):
        check_apply(n, op, dtype, k)


@pytest.mark.parametrize("dtype", DTYPES)
def test_flip_wider_than_a_tile_falls_back(dtype):
    """X on all 20 qubits: the flip has more high qubits than a tile holds -> gather kernel."""
    n = 20
    op = mq.QubitOperator(" ".join(f"X{q}" for q in range(n)), 0.7) + mq.QubitOperator("Z3 Y17", -0.4)
    check_apply(n, op, dtype, 3)


@pytest.mark.parametrize("dtype", DTYPES)
def test_wide_tile_kept_when_a_flip_needs_it(dtype):
    """A flip over 9 qubits above the low run fits the default tile width (14 / 13 qubits for
    complex64 / complex128) but not one qubit less: the plan keeps the wider tiles.  The ring terms
    beside it (which alone would take the narrower tiles, test_apply_hamiltonian_ring_multi_pass)
    share the plan."""
    n = 20
    low = 5 if dtype == mq.complex64 else 4
    wide = mq.QubitOperator(" ".join(f"X{q}" for q in range(low, low + 9)), 0.6)
    op = wide + mq.QubitOperator(f"Y{low + 2} Z{low + 8}", -0.3) + ring_hamiltonian(n, 13)
    check_apply(n, op, dtype, 13)


@pytest.mark.parametrize("dtype", DTYPES)
def test_tile_engine_matches_gather_engine(dtype):
    n = 17
    op = random_hamiltonian(n, 30, 5, max_weight=5) + ring_hamiltonian(n, 5)
    a = check_apply(n, op, dtype, 11)
    with options(ham_tile_engine=0):
        b = check_apply(n, op, dtype, 11)
    assert np.max(np.abs(a - b)) < tol_for(op, dtype)


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("n,batch", [(15, 3), (19, 2)])
def test_grad_with_tile_hamiltonian(n, batch, dtype):
    """The batched adjoint gradient starts from lambda = H psi and f = <psi|H|psi> of the tile engine."""
    circ = hea(n, 2, encoder=True)
    ops = [ring_hamiltonian(n, n), random_hamiltonian(n, 12, n, max_weight=4)]
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    rng = np.random.default_rng(n)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g_enc, g_ans = sim.get_expectation_with_grad(hams, circ)(enc, ans)
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    f_ref, g_ref = cpu.grad_batch(n, np.complex128, ir.lower(circ, names),
                                  [mq.Hamiltonian(o, mq.complex128).terms_ir() for o in ops], enc, ans, batch_threads=4)
    g = np.concatenate([g_enc, g_ans], axis=2)
    for h, o in enumerate(ops):
        assert np.max(np.abs(f[:, h] - f_ref[:, h])) < tol_for(o, dtype)
        assert_g(g[:, h], g_ref[:, h], dtype)


@pytest.mark.parametrize("dtype", DTYPES)
def test_grad_recomputes_forward_when_copy_exceeds_budget(dtype):
    """Several Hamiltonians with an HBM budget below three states per sample: the forward state is
    recomputed per Hamiltonian instead of kept; results equal the unconstrained run."""
    n, batch = 16, 3
    circ = hea(n, 3, encoder=True)
    ops = [ring_hamiltonian(n, 1), random_hamiltonian(n, 10, 2, max_weight=3), ring_hamiltonian(n, 3)]
    hams = [mq.Hamiltonian(o, dtype) for o in ops]
    rng = np.random.default_rng(5)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f0, ge0, ga0 = sim.get_expectation_with_grad(hams, circ)(enc, ans)
    state = (1 << n) * np.dtype(dtype).itemsize
    before = _native.get_option("grad_recomputes")
    with options(hbm_budget_bytes=int(2.5 * state)):
        f1, ge1, ga1 = sim.get_expectation_with_grad(hams, circ)(enc, ans)
    # the budget holds one sample (two states) per chunk: every sample recomputes for ops 2 and 3
    assert _native.get_option("grad_recomputes") == before + batch * (len(ops) - 1)
    tol = 1e-12 if dtype == mq.complex128 else 1e-6
    assert np.max(np.abs(f1 - f0)) < tol and np.max(np.abs(ge1 - ge0)) < tol and np.max(np.abs(ga1 - ga0)) < tol
