"""The seeded benchmark workloads, shared by bench.py and the parity tests that check them at the
size they are timed.

* ``ring_x_hamiltonian``: sum_i w_i Z_i Z_{i+1 mod n} + sum_i v_i X_i, w, v ~ U(-1, 1) (the "n ZZ + n X"
  terms read as a ring).
* ``batched_adjoint``: encoder RX(x_i) on every qubit (x ~ N(0, 1), shape [batch, n]) followed by an
  HEA of ``layers`` layers (RY, RZ on every qubit, CNOT ladder i -> i+1; parameters ~ U[0, 2 pi)),
  with ``ring_x_hamiltonian`` - the 20-qubit batch-1024 gradient.
* ``hea_open_chain``: the CPU-runnable single-vector case - HEA of ``layers`` layers on n qubits,
  H = sum_i Z_i Z_{i+1} (open chain) + sum_i 0.5 X_i, parameters ~ U[0, 2 pi) from ``seed``.
* ``large_hea``: HEA of ``layers`` layers with ``n_terms`` random Pauli strings of weight 2-4 and
  N(0, 1) coefficients - the 32-qubit / 31-qubit single-instance gradients and their 24-qubit
  analog.
"""
import numpy as np

from ..core.operators import QubitOperator
from .builders import encoder_rx, hardware_efficient_ansatz, random_pauli_hamiltonian


def ring_x_hamiltonian(n, rng):
    op = QubitOperator("", 0)
    w = rng.uniform(-1, 1, size=n)
    v = rng.uniform(-1, 1, size=n)
    for i in range(n):
        op += QubitOperator(f"Z{i} Z{(i + 1) % n}", float(w[i]))
        op += QubitOperator(f"X{i}", float(v[i]))
    return op


def batched_adjoint(n=20, layers=8, batch=1024, seed=2024):
    """(circuit, QubitOperator, encoder rows [batch, n], ansatz vector)."""
    rng = np.random.default_rng(seed)
    circ = encoder_rx(n) + hardware_efficient_ansatz(n, layers)
    op = ring_x_hamiltonian(n, rng)
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    return circ, op, enc, ans


def hea_open_chain(n=16, layers=10, seed=0):
    """(circuit, QubitOperator, parameter vector)."""
    circ = hardware_efficient_ansatz(n, layers)
    op = QubitOperator("", 0)
    for i in range(n - 1):
        op += QubitOperator(f"Z{i} Z{i + 1}")
    for i in range(n):
        op += 0.5 * QubitOperator(f"X{i}")
    vals = np.random.default_rng(seed).uniform(0, 2 * np.pi, size=len(circ.params_name))
    return circ, op, vals


def large_hea(n, layers=4, n_terms=64, seed=2024):
    """(circuit, QubitOperator, parameter vector)."""
    rng = np.random.default_rng(seed)
    circ = hardware_efficient_ansatz(n, layers)
    op = random_pauli_hamiltonian(n, n_terms, seed)
    vals = rng.uniform(0, 2 * np.pi, size=len(circ.params_name))
    return circ, op, vals
