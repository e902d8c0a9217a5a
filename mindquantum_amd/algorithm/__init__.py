"""Circuit builders used by the benchmarks and tests (mindquantum.algorithm.nisq subset)."""
from .builders import (hardware_efficient_ansatz, layered_random_circuit, max_cut_ansatz, max_cut_hamiltonian,
                       qaoa_graph, qft, random_pauli_hamiltonian, random_regular_graph,
                       uneven_qaoa_batch)
from .workloads import batched_adjoint, hea_open_chain, large_hea, ring_x_hamiltonian

__all__ = ["hardware_efficient_ansatz", "layered_random_circuit", "max_cut_ansatz", "max_cut_hamiltonian",
           "qaoa_graph", "qft", "random_pauli_hamiltonian", "random_regular_graph", "uneven_qaoa_batch",
           "batched_adjoint", "hea_open_chain", "large_hea", "ring_x_hamiltonian"]
