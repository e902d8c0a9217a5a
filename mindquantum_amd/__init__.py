"""mindquantum_amd: MindQuantum's state-vector simulator path, native on MI355X (gfx950).

``Simulator('mqvector_amd', n)`` keeps MindQuantum's Simulator API (apply_circuit, get_qs,
get_expectation, get_expectation_with_grad); every gate, expectation and adjoint gradient runs
in the HIP library ``libmqamd.so`` (C ABI in include/mqamd.h).
"""
from . import dtype
from .core import (CNOT, FSIM, H, I, ISWAP, RX, RY, RZ, SWAP, SX, U3, UN, BarrierGate, Circuit, FSim, Givens,
                   GlobalPhase, GroupedPauli, Hamiltonian, Measure, MeasureResult, ParameterResolver, PhaseShift, Power, QubitOperator, Rn,
                   RotPauliString, Rxx, Rxy, Rxz, Ryy, Ryz, Rzz, S, SWAPalpha, T, UnivMathGate, X, Y, Z)
from .core import gates
from .core.gates import gene_univ_parameterized_gate, gene_univ_two_params_gate
from .dtype import complex64, complex128
from .simulator import Simulator, expectation_with_grad_batch, inner_product

__version__ = "0.1.0"
