"""The allocation-free expectation entry point (mqa_sim_expectation_noalloc) against the entry point
it varies, mqa_sim_expectation with no circuit, both called through the C ABI: two calls in a row
(the second on a smaller state and a different Hamiltonian, so the shared device scratch holds the
first call's contents), in both precisions; the same kernels make the results identical.  The values
are also checked against the complex128 oracle's <psi|H|psi> (bars of tests/bars.py), and a
Hamiltonian beyond the scratch's term capacity is rejected with ValueError.
"""
import ctypes

import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.algorithm import hardware_efficient_ansatz, random_pauli_hamiltonian
from oracle import cpu
from tests.bars import assert_f, coeff_sum

pytestmark = pytest.mark.gpu


def both(sim, ham):
    lib = _native.lib()
    a = np.zeros(2)
    b = np.zeros(2)
    _native.check(lib.mqa_sim_expectation_noalloc(sim.backend._h, ham.native(), _native.ptr(a)))
    _native.check(lib.mqa_sim_expectation(sim.backend._h, ham.native(), None, None, 0, _native.ptr(b)))
    return complex(a[0], a[1]), complex(b[0], b[1])


@pytest.mark.parametrize("dtype", [mq.complex64, mq.complex128])
def test_noalloc_matches_expectation_on_two_calls(dtype):
    for n, n_terms, seed in ((20, 40, 1), (11, 12, 2)):
        circ = hardware_efficient_ansatz(n, 2)
        vals = np.random.default_rng(seed).uniform(0, 2 * np.pi, size=len(circ.params_name))
        op = random_pauli_hamiltonian(n, n_terms, seed, min_weight=1, max_weight=4) + mq.QubitOperator("", 0.75)
        ham = mq.Hamiltonian(op, dtype)
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        sim.apply_circuit(circ, vals)
        got, want = both(sim, ham)
        assert got == want, f"{n}q: {got} vs {want}"
        psi = sim.get_qs().astype(np.complex128)
        ref = cpu.expectation(psi, psi, mq.Hamiltonian(op, mq.complex128).terms_ir())
        assert_f(got, ref, dtype, coeff_sum(op))


def test_noalloc_rejects_hamiltonians_beyond_its_scratch():
    terms = np.zeros(70000, dtype=_native.TERM_DTYPE)
    terms["z_mask"] = np.arange(1, 70001, dtype=np.uint64)  # 70000 distinct Z strings on 17 qubits
    terms["coeff"] = 1e-3
    h = ctypes.c_void_p()
    _native.check(_native.lib().mqa_ham_create(_native.ptr(terms), len(terms), ctypes.byref(h)))
    try:
        sim = mq.Simulator("mqvector_amd", 17, dtype=mq.complex64)
        out = np.zeros(2)
        st = _native.lib().mqa_sim_expectation_noalloc(sim.backend._h, h, _native.ptr(out))
        assert st == _native.MQA_ERR_ARG
        assert "at most" in _native.lib().mqa_last_error().decode()
    finally:
        _native.lib().mqa_ham_destroy(h)
