"""Seeded circuit / Hamiltonian builders for the benchmark workloads.

* hardware_efficient_ansatz: layers of RY, RZ on every qubit and a CNOT ladder i -> i+1
  (algorithm/nisq/chem/hardware_efficient_ansatz.py:45,111-119 builds the same layer shape).
* max_cut_ansatz / max_cut_hamiltonian: H on every node, then p x [Rzz(gamma_l) on every edge,
  RX(beta_l) on every node]; cost sum_edges (1 - Z_u Z_v)/2 (qaoa/max_cut_ansatz.py:119-141,
  :153, :229-232).
* layered_random_circuit: the 28-qubit random-circuit workload - per layer H / RX / RY / RZ on
  every qubit, then CZ or CNOT on a seeded random perfect matching.
"""
import numpy as np

from ..core import gates as G
from ..core.circuit import UN, Circuit
from ..core.operators import QubitOperator


def hardware_efficient_ansatz(n_qubits, layers, prefix=""):
    circ = Circuit()
    for layer in range(layers):
        for q in range(n_qubits):
            circ += G.RY(f"{prefix}ry{layer}_{q}").on(q)
        for q in range(n_qubits):
            circ += G.RZ(f"{prefix}rz{layer}_{q}").on(q)
        for q in range(n_qubits - 1):
            circ += G.X.on(q + 1, q)
    return circ


def encoder_rx(n_qubits, prefix="x"):
    circ = Circuit([G.RX(f"{prefix}{q}").on(q) for q in range(n_qubits)])
    return circ.as_encoder()


def random_regular_graph(n, degree, seed):
    """Seeded random d-regular graph by the configuration model with restarts (n * d even)."""
    if (n * degree) % 2:
        raise ValueError("n * degree must be even")
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        stubs = np.repeat(np.arange(n), degree)
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for a, b in zip(stubs[0::2], stubs[1::2]):
            a, b = int(a), int(b)
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)
    raise RuntimeError("failed to sample a regular graph")


def qaoa_graph(n, seed):
    """Seeded 3-regular graph on n nodes; for odd n (no 3-regular graph exists) the closest one:
    every node of degree 3 except node n - 1, of degree 2."""
    if n % 2 == 0:
        return random_regular_graph(n, 3, seed)
    rng = np.random.default_rng(seed)
    for _ in range(10000):
        stubs = np.concatenate([np.repeat(np.arange(n - 1), 3), [n - 1, n - 1]])
        rng.shuffle(stubs)
        edges = set()
        ok = True
        for a, b in zip(stubs[0::2], stubs[1::2]):
            a, b = int(a), int(b)
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)
    raise RuntimeError("failed to sample a graph")


def uneven_qaoa_batch(count, seed, n_min=12, n_max=24, depth=4):
    """The uneven MaxCut batch: `count` circuits with qubit counts uniform in [n_min, n_max],
    each a QAOA ansatz of `depth` layers on its own seeded graph (qaoa_graph) with angles
    gamma, beta ~ U(0, pi).  Returns (circuits, hamiltonians (QubitOperator), qubit counts,
    parameter vectors in each circuit's params_name order)."""
    rng = np.random.default_rng(seed)
    circs, ops, nqs, params = [], [], [], []
    for i in range(count):
        n = int(rng.integers(n_min, n_max + 1))
        graph = qaoa_graph(n, int(rng.integers(1 << 31)))
        c = max_cut_ansatz(graph, depth)
        circs.append(c)
        ops.append(max_cut_hamiltonian(graph))
        nqs.append(n)
        params.append(rng.uniform(0, np.pi, size=len(c.params_name)))
    return circs, ops, nqs, params


def max_cut_ansatz(graph, depth):
    nodes = sorted({q for e in graph for q in e})
    circ = UN(G.H, nodes)
    for p in range(depth):
        for u, v in graph:
            circ += G.Rzz(f"gamma_{p}").on([u, v])
        for q in nodes:
            circ += G.RX(f"beta_{p}").on(q)
    return circ


def max_cut_hamiltonian(graph):
    op = QubitOperator("", 0)
    for u, v in graph:
        op += (QubitOperator("") - QubitOperator(f"Z{u} Z{v}")) / 2
    return op


def layered_random_circuit(n_qubits, depth, seed):
    """Per layer: H / RX / RY / RZ (uniform angle) on every qubit, then CZ or CNOT on a random
    perfect matching (the first qubit of a pair is the control)."""
    rng = np.random.default_rng(seed)
    circ = Circuit()
    for _ in range(depth):
        for q in range(n_qubits):
            k = int(rng.integers(4))
            if k == 0:
                circ += G.H.on(q)
            else:
                th = float(rng.uniform(0, 2 * np.pi))
                circ += (G.RX, G.RY, G.RZ)[k - 1](th).on(q)
        perm = rng.permutation(n_qubits)
        for a, b in zip(perm[0::2], perm[1::2]):
            if rng.integers(2):
                circ += G.Z.on(int(b), int(a))
            else:
                circ += G.X.on(int(b), int(a))
    return circ


def qft(qubits):
    """Quantum Fourier transform on `qubits` (algorithm/library/quantum_fourier.py): per qubit i, H on
    it and PhaseShift(2 pi / 2^k) controlled by each later qubit (k = 2, 3, ...), then the qubit order
    reversed by SWAPs."""
    qubits = list(qubits)
    circ = Circuit()
    for i, q in enumerate(qubits):
        circ += G.H.on(q)
        for k, c in enumerate(qubits[i + 1:]):
            circ += G.PhaseShift(2 * np.pi / 2 ** (k + 2)).on(q, c)
    for j in range(len(qubits) // 2):
        circ += G.SWAP.on([qubits[j], qubits[len(qubits) - 1 - j]])
    return circ


def random_pauli_hamiltonian(n_qubits, n_terms, seed, min_weight=2, max_weight=4):
    rng = np.random.default_rng(seed)
    op = QubitOperator("", 0)
    for _ in range(n_terms):
        w = int(rng.integers(min_weight, max_weight + 1))
        qs = rng.choice(n_qubits, size=w, replace=False)
        op += QubitOperator(" ".join(f"{'XYZ'[rng.integers(3)]}{q}" for q in qs), float(rng.normal()))
    return op
