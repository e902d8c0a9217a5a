"""Phase-free plans of the hermitian gradient (option ``phase_free``): uncontrolled RX / RY / RZ /
GP drop their global phase in the forward sweep, the back-sweep and the checkpoints alike, and
the RZ / GP taps use that <lambda|psi> = <psi|H|psi> is real.  Gradients must match the oracle
with and without them, in the interpreted and in the compiled passes.

Narrow tiles (7 of 12 qubits local) make most controls of parameterised gates non-local, so a
tapped op whose control is off for a tile sits between other tapped ops of its register group
(the interpreted pass must still advance its tap slot).  Bars of tests/bars.py against the
complex128 oracle.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.core import gates as G
from tests.bars import assert_f, coeff_sum
from tests.conftest import hea, params_for, random_circuit, random_hamiltonian
from tests.test_gpu_parity import DTYPES, assert_grad, oracle_grad

pytestmark = pytest.mark.gpu

KEYS = ("jit", "jit_min_amps", "phase_free", "checkpoints")


@pytest.fixture
def options():
    saved = {k: _native.get_option(k) for k in KEYS}
    yield
    for k, v in saved.items():
        _native.set_option(k, v)


def set_mode(compiled, phase_free):
    _native.set_option("jit", 2 if compiled else 0)
    _native.set_option("jit_min_amps", 0)
    _native.set_option("phase_free", int(phase_free))


def narrow(circ, dtype, k):
    code = mq.dtype.native_code(dtype)
    circ._mqa_tiles = {(code, False): k, (code, True): k}
    return circ


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compiled", [False, True])
@pytest.mark.parametrize("phase_free", [False, True])
def test_random_circuit_non_local_controls(options, dtype, compiled, phase_free):
    n = 12
    set_mode(compiled, phase_free)
    circ = narrow(random_circuit(n, 120, 41, ctrl_prob=0.4), dtype, 7)
    op = random_hamiltonian(n, 12, 8)
    names, vals = params_for(circ, 5)
    before = _native.get_option("jit_launches")
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    assert _native.get_option("phase_free_used") == int(phase_free)
    assert names == list(circ.ansatz_params_name)
    f_ref, g_ref = oracle_grad(circ, n, [op], np.zeros((1, 0)), vals)
    assert_f(f[0, 0], f_ref[0, 0], dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref[0, 0], dtype)
    assert (_native.get_option("jit_launches") > before) == compiled


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compiled", [False, True])
def test_hea_batch_with_global_phases_and_checkpoints(options, dtype, compiled):
    """Encoder RX, HEA RY / RZ over every angle quadrant, GP gates (zero taps), checkpoints."""
    n, batch = 14, 3
    set_mode(compiled, True)
    _native.set_option("checkpoints", 3)
    circ = hea(n, 3, encoder=True)
    circ += G.GlobalPhase("gp0").on(2)
    circ += G.GlobalPhase({"gp1": 0.7, "a0": -0.3}).on(9)
    circ += G.RZ("rz_c").on(5, 11)  # controlled: stays exact
    circ = narrow(circ, dtype, 10)
    ops = [mq.QubitOperator("Z0 Z1") + 0.5 * mq.QubitOperator("X13"), random_hamiltonian(n, 8, 3)]
    rng = np.random.default_rng(7)
    enc = rng.uniform(-2 * np.pi, 2 * np.pi, size=(batch, len(circ.encoder_params_name)))
    ans = rng.uniform(-2 * np.pi, 2 * np.pi, size=len(circ.ansatz_params_name))
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g_enc, g_ans = sim.get_expectation_with_grad([mq.Hamiltonian(o, dtype) for o in ops], circ)(enc, ans)
    assert _native.get_option("checkpoints_used") > 0
    f_ref, g_ref = oracle_grad(circ, n, ops, enc, ans)
    for h, o in enumerate(ops):
        assert_f(f[:, h], f_ref[:, h], dtype, coeff_sum(o))
    assert_grad(np.concatenate([g_enc, g_ans], axis=2), g_ref, dtype)


def test_phase_free_matches_exact_plans(options):
    """The same gradient with both plan forms, complex128, to rounding."""
    n, batch = 16, 2
    circ = hea(n, 4, encoder=True)
    op = random_hamiltonian(n, 10, 11)
    rng = np.random.default_rng(2)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 4 * np.pi, size=len(circ.ansatz_params_name))
    out = []
    for pf in (False, True):
        set_mode(True, pf)
        sim = mq.Simulator("mqvector_amd", n, dtype=mq.complex128)
        out.append(sim.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex128), circ)(enc, ans))
    for a, b in zip(out[0], out[1]):
        assert np.max(np.abs(a - b)) < 1e-11


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compiled", [False, True])
def test_apply_circuit_restores_the_dropped_phase(options, dtype, compiled):
    """apply_circuit runs the phase-free plan and multiplies the dropped global phase back in:
    the state (its phase included) matches the oracle and the exact plan, with half-angle cosines
    of both signs, global-phase gates, and controlled rotations (which stay exact)."""
    from tests.test_gpu_parity import assert_state, oracle_state

    n = 11
    rng = np.random.default_rng(12)
    circ = mq.Circuit()
    for q in range(n):
        circ.h(q)
    for k in range(60):
        q = int(rng.integers(n))
        kind = k % 5
        th = float(rng.uniform(-3 * np.pi, 3 * np.pi))
        if kind == 0:
            circ += G.RX(th).on(q)
        elif kind == 1:
            circ += G.RY(f"p{k}").on(q)
        elif kind == 2:
            circ += G.RZ(th).on(q)
        elif kind == 3:
            circ += G.GlobalPhase(f"p{k}").on(q)
        else:
            circ += G.RY(th).on(q, (q + 1) % n)
        if k % 7 == 0:
            circ += G.X.on((q + 2) % n, q)
    names = list(circ.params_name)
    vals = dict(zip(names, rng.uniform(-3 * np.pi, 3 * np.pi, size=len(names))))
    states = []
    for pf in (True, False):
        set_mode(compiled, pf)
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        sim.apply_circuit(circ, vals)
        states.append(sim.get_qs())
    assert_state(states[0], oracle_state(circ, n, vals, dtype), dtype)
    tol = 1e-12 if dtype == mq.complex128 else 2e-6
    assert np.max(np.abs(states[0] - states[1])) < tol


@pytest.mark.parametrize("dtype", DTYPES)
@pytest.mark.parametrize("compiled", [False, True])
def test_rzz_phase_free_gradient_and_state(options, dtype, compiled):
    """Uncontrolled Rzz runs as a phase on the odd-parity pair: gradients (with checkpoints) and
    apply_circuit states (dropped phase restored) match the oracle; controlled Rzz stays exact."""
    from tests.test_gpu_parity import assert_state, oracle_state

    n = 12
    rng = np.random.default_rng(21)
    circ = mq.Circuit()
    for q in range(n):
        circ += G.RX(f"x{q}").on(q)
    for k in range(40):
        a, b = (int(v) for v in rng.choice(n, 2, replace=False))
        if k % 6 == 5:
            circ += G.Rzz(f"z{k}").on([a, b], int((max(a, b) + 1) % n) if (max(a, b) + 1) % n not in (a, b) else None)
        else:
            circ += G.Rzz(f"z{k}").on([a, b])
        circ += G.RX(f"m{k}").on(int(rng.integers(n)))
    circ = narrow(circ, dtype, 8)
    op = random_hamiltonian(n, 10, 3)
    names = list(circ.params_name)
    vals = rng.uniform(-2 * np.pi, 2 * np.pi, size=len(names))
    set_mode(compiled, True)
    _native.set_option("checkpoints", 2)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vals)
    assert _native.get_option("phase_free_used") == 1
    f_ref, g_ref = oracle_grad(circ, n, [op], np.zeros((1, 0)), vals)
    assert_f(f[0, 0], f_ref[0, 0], dtype, coeff_sum(op))
    assert_grad(g[0, 0].real, g_ref[0, 0], dtype)
    sim2 = mq.Simulator("mqvector_amd", n, dtype=dtype)
    pr = dict(zip(names, vals))
    sim2.apply_circuit(circ, pr)
    assert_state(sim2.get_qs(), oracle_state(circ, n, pr, dtype), dtype)
