"""Checks at the full sizes the benchmark times, through properties the mathematics fixes (no oracle
can run there in seconds).

* 28 qubits, complex64 and complex128: the layered random circuit (H / RX / RY / RZ on every qubit,
  CZ / CNOT on a random perfect matching) for 20 layers followed by its hermitian returns |0...0>:
  1 - |<0|psi>| <= 1e-5 (complex64) / 1e-12 (complex128), the round trip of the reference's
  test_simulator.py:125-167.  The full depth-200 circuit keeps the norm to the same bars.
* 32 qubits complex64 and 31 qubits complex128 (34.4 GB states), with 4 checkpoints: RY(a_q) RZ(b_q)
  on every qubit, a CNOT ladder and its inverse, RY(c_q) RZ(d_q) on every qubit, and
  H = sum_i w_i Z_i Z_{i+1} + sum_i v_i X_i.  The ladders cancel, so the state is the product of the
  one-qubit states RZ(d) RY(c) RZ(b) RY(a)|0> and f and every derivative are closed forms in the
  one-qubit Bloch components (computed here from 2 x 2 matrices in complex128).  Bars: 1e-5 (complex64)
  / 1e-10 (complex128) times sum |c| for f, relative to the largest |g| for the gradient.  This runs
  the tile passes, the Hamiltonian engine, the reductions and the checkpoint pointer switch with
  amplitude offsets past 2^32.
"""
import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native
from mindquantum_amd.algorithm import layered_random_circuit
from mindquantum_amd.core import gates as G
from tests.bars import assert_f, assert_g

pytestmark = pytest.mark.gpu

DTYPES = [mq.complex64, mq.complex128]


def fidelity_bar(dtype):
    return 1e-12 if dtype == mq.complex128 else 1e-5


@pytest.mark.parametrize("dtype", DTYPES)
def test_28q_round_trip_and_norm(dtype):
    n = 28
    circ = layered_random_circuit(n, 20, 2024)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
    sim.apply_circuit(circ + circ.hermitian())
    zero = mq.Simulator("mqvector_amd", n, dtype=dtype)
    overlap = abs(mq.inner_product(zero, sim))
    print(f"{np.dtype(dtype).name}: 1 - |<0|U^dag U|0>| = {1 - overlap:.3e}")
    assert abs(1 - overlap) <= fidelity_bar(dtype)
    del zero
    full = layered_random_circuit(n, 200, 2024)
    sim.reset()
    sim.apply_circuit(full)
    norm = mq.inner_product(sim, sim).real
    print(f"{np.dtype(dtype).name}: | ||psi||^2 - 1 | after depth 200 = {abs(norm - 1):.3e}")
    assert abs(norm - 1) <= fidelity_bar(dtype)


def _rot(kind, th):
    c, s = np.cos(th / 2), np.sin(th / 2)
    if kind == "RY":
        return np.array([[c, -s], [s, c]], dtype=np.complex128)
    return np.array([[c - 1j * s, 0], [0, c + 1j * s]], dtype=np.complex128)


def _drot(kind, th):
    c, s = np.cos(th / 2), np.sin(th / 2)
    if kind == "RY":
        return 0.5 * np.array([[-s, -c], [c, -s]], dtype=np.complex128)
    return 0.5 * np.array([[-s - 1j * c, 0], [0, -s + 1j * c]], dtype=np.complex128)


def product_circuit(n):
    circ = mq.Circuit()
    for q in range(n):
        circ += G.RY(f"a{q}").on(q)
        circ += G.RZ(f"b{q}").on(q)
    for q in range(n - 1):
        circ += G.X.on(q + 1, q)
    for q in reversed(range(n - 1)):
        circ += G.X.on(q + 1, q)
    for q in range(n):
        circ += G.RY(f"c{q}").on(q)
        circ += G.RZ(f"d{q}").on(q)
    return circ


def closed_form(n, vals, w, v):
    """f and df/dparam for product_circuit with H = sum w_i Z_i Z_{i+1} + sum v_i X_i."""
    X = np.array([[0, 1], [1, 0]], dtype=np.complex128)
    Z = np.diag([1.0, -1.0]).astype(np.complex128)
    seq = ("RY", "RZ", "RY", "RZ")
    z, x = np.zeros(n), np.zeros(n)
    dz, dx = np.zeros((n, 4)), np.zeros((n, 4))
    for q in range(n):
        th = vals[q]
        mats = [_rot(k, t) for k, t in zip(seq, th)]
        phi = mats[3] @ mats[2] @ mats[1] @ mats[0] @ np.array([1, 0], dtype=np.complex128)
        z[q] = np.vdot(phi, Z @ phi).real
        x[q] = np.vdot(phi, X @ phi).real
        for j in range(4):
            ms = list(mats)
            ms[j] = _drot(seq[j], th[j])
            dphi = ms[3] @ ms[2] @ ms[1] @ ms[0] @ np.array([1, 0], dtype=np.complex128)
            dz[q, j] = 2 * np.vdot(phi, Z @ dphi).real
            dx[q, j] = 2 * np.vdot(phi, X @ dphi).real
    f = sum(w[i] * z[i] * z[i + 1] for i in range(n - 1)) + float(np.dot(v, x))
    g = np.zeros((n, 4))
    for q in range(n):
        g[q] += v[q] * dx[q]
        if q > 0:
            g[q] += w[q - 1] * z[q - 1] * dz[q]
        if q < n - 1:
            g[q] += w[q] * z[q + 1] * dz[q]
    return f, g


@pytest.mark.parametrize("n,dtype", [(32, mq.complex64), (31, mq.complex128)])
def test_large_product_state_gradient_with_checkpoints(n, dtype):
    rng = np.random.default_rng(n)
    circ = product_circuit(n)
    w = rng.uniform(-1, 1, size=n - 1)
    v = rng.uniform(-1, 1, size=n)
    op = mq.QubitOperator("", 0)
    for i in range(n - 1):
        op += mq.QubitOperator(f"Z{i} Z{i + 1}", float(w[i]))
    for i in range(n):
        op += mq.QubitOperator(f"X{i}", float(v[i]))
    th = rng.uniform(0, 2 * np.pi, size=(n, 4))
    values = {}
    for q in range(n):
        for j, p in enumerate("abcd"):
            values[f"{p}{q}"] = th[q, j]
    names = list(circ.params_name)
    vec = np.array([values[k] for k in names])
    f_ref, g_ref = closed_form(n, th, w, v)
    g_ref_vec = np.array([g_ref[int(k[1:]), "abcd".index(k[0])] for k in names])

    saved = {k: _native.get_option(k) for k in ("checkpoints", "jit")}
    _native.set_option("checkpoints", 4)
    _native.set_option("jit", 2)
    try:
        sim = mq.Simulator("mqvector_amd", n, dtype=dtype)
        f, g = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)(vec)
        used = _native.get_option("checkpoints_used")
    finally:
        for k, val in saved.items():
            _native.set_option(k, val)
    csum = float(np.sum(np.abs(w)) + np.sum(np.abs(v)))
    err_f = assert_f(f[0, 0], f_ref, dtype, csum)
    err_g = assert_g(g[0, 0].real, g_ref_vec, dtype)
    print(f"{n}q {np.dtype(dtype).name}: checkpoints {used}, |f - f_ref| = {err_f:.2e}, "
          f"max|g - g_ref| = {err_g:.2e} (max|g_ref| = {np.max(np.abs(g_ref_vec)):.3f})")
    assert used == 4
