"""Dense gate fusion: runs of gates merged on the host into dense blocks of at most k qubits.

A fused circuit runs in one of two ways:

* ``FusedCircuit.apply(sim)``: every block is one streaming sweep over the state
  (``mqa_sim_apply_dense_blocks``; complex64 5-qubit blocks on the FP32 matrix cores), the GPU form
  of the reference's dense apply ``ApplyNQubitsMatrix``
  (ccsrc/lib/simulator/vector/detail/cpu_common/cpu_vector_core_matrix_gate.cpp:28-83) with one
  sweep per fused block instead of one per gate.
* ``FusedCircuit.as_circuit()``: the blocks as ``UnivMathGate`` s (k <= 4) for the tile-pass
  scheduler, which packs several of them into each pass over HBM like any other gates.

The fuser is greedy and keeps the circuit's order: each qubit belongs to at most one open block; a
gate joins the union of the open blocks on its qubits when that union has at most k qubits, and
otherwise those blocks are closed (emitted) and the gate opens a new one.  A closed block is emitted
before any later gate on its qubits, and the blocks still open act on other qubits, so the product
of the emitted blocks equals the circuit.  The reference simulator has no fusion pass; this module is
an addition of this backend.
"""
import numpy as np

from . import _native
from .core import gates as G
from .core.circuit import Circuit
from .core.measurement import Measure

MAX_FUSED_QUBITS = 5


class DenseBlock:
    """A dense block: ``matrix`` (2^k x 2^k complex128) on ``qubits`` (ascending; matrix index bit b
    <-> qubits[b]) equal to the product of ``gates`` (in circuit order)."""

    def __init__(self, qubits, matrix, gates):
        self.qubits = tuple(int(q) for q in qubits)
        self.matrix = matrix
        self.gates = gates

    def __repr__(self):
        return f"DenseBlock(qubits={self.qubits}, gates={len(self.gates)})"


def _gate_matrix(g, pr):
    if g.param_resolvers():
        if pr is None and g.parameterized:
            raise ValueError(f"fusion of the parameterised gate {g!r} needs a parameter resolver")
        return np.asarray(g.matrix(pr if pr is not None else {}), dtype=np.complex128)
    return np.asarray(g.matrix(), dtype=np.complex128)


def _apply(u, m, obj_pos, ctrl_pos):
    """u <- (m on block positions obj_pos, controlled by ctrl_pos) @ u, u of shape (D, D)."""
    d = u.shape[0]
    idx = np.arange(d)
    cmask = sum(1 << p for p in ctrl_pos)
    omask = sum(1 << p for p in obj_pos)
    bases = idx[((idx & cmask) == cmask) & ((idx & omask) == 0)]
    sub = np.array([sum(((r >> b) & 1) << p for b, p in enumerate(obj_pos)) for r in range(1 << len(obj_pos))])
    rows = bases[:, None] | sub[None, :]
    u[rows, :] = np.einsum("rc,bcj->brj", m, u[rows, :])
    return u


def _block_matrix(qubits, gates, pr):
    pos = {q: p for p, q in enumerate(qubits)}
    u = np.eye(1 << len(qubits), dtype=np.complex128)
    for g in gates:
        _apply(u, _gate_matrix(g, pr), [pos[q] for q in g.obj_qubits], [pos[q] for q in g.ctrl_qubits])
    return u


def _primitives(circ):
    for g in circ:
        if isinstance(g, Measure):
            raise ValueError("fusion does not apply to circuits with measurements")
        if isinstance(g, G.BarrierGate) or not g.obj_qubits:
            continue
        yield g


def fuse(circ, max_qubits, pr=None):
    """Fuse ``circ`` into dense blocks of at most ``max_qubits`` (1..5) qubits.  Parameterised gates
    take their values from ``pr`` (a dict or ParameterResolver); a gate on more qubits than
    ``max_qubits`` is a block of its own (at most 5 qubits).  Returns a FusedCircuit."""
    k = int(max_qubits)
    if not 1 <= k <= MAX_FUSED_QUBITS:
        raise ValueError(f"max_qubits must be in 1..{MAX_FUSED_QUBITS}")
    open_blocks = {}  # block id -> (set of qubits, list of gates)
    owner = {}  # qubit -> open block id
    closed = []
    next_id = 0

    def close(bid):
        qs, gs = open_blocks.pop(bid)
        for q in qs:
            del owner[q]
        closed.append((sorted(qs), gs))

    for g in _primitives(circ):
        qs = set(g.obj_qubits) | set(g.ctrl_qubits)
        if len(qs) > MAX_FUSED_QUBITS:
            raise ValueError(f"{g!r} acts on more than {MAX_FUSED_QUBITS} qubits")
        touching = sorted({owner[q] for q in qs if q in owner})
        union = set(qs)
        for b in touching:
            union |= open_blocks[b][0]
        gates = []
        if len(union) <= k:  # merge: the open blocks act on disjoint qubits, so their gates commute
            for b in touching:
                qsb, gsb = open_blocks.pop(b)
                for q in qsb:
                    del owner[q]
                gates += gsb
        else:
            for b in touching:
                close(b)
            union = set(qs)
        gates.append(g)
        open_blocks[next_id] = (union, gates)
        for q in union:
            owner[q] = next_id
        next_id += 1
    for b in sorted(open_blocks):
        close(b)
    blocks = [DenseBlock(qs, _block_matrix(qs, gs, pr), gs) for qs, gs in closed]
    return FusedCircuit(blocks, k)


class FusedCircuit:
    """The dense blocks of a fused circuit, in application order."""

    def __init__(self, blocks, max_qubits):
        self.blocks = list(blocks)
        self.max_qubits = max_qubits

    def __len__(self):
        return len(self.blocks)

    @property
    def n_gates(self):
        return sum(len(b.gates) for b in self.blocks)

    def histogram(self):
        """{block width: number of blocks}."""
        h = {}
        for b in self.blocks:
            h[len(b.qubits)] = h.get(len(b.qubits), 0) + 1
        return dict(sorted(h.items()))

    def arrays(self):
        """(block_qubits int32[n], qubits int32[sum k], matrices float64[sum 2 * 4^k]) of the C ABI."""
        nq = np.array([len(b.qubits) for b in self.blocks], dtype=np.int32)
        qs = np.array([q for b in self.blocks for q in b.qubits], dtype=np.int32)
        mats = np.concatenate([np.ascontiguousarray(b.matrix).view(np.float64).ravel() for b in self.blocks]) \
            if self.blocks else np.zeros(0)
        return nq, qs, np.ascontiguousarray(mats, dtype=np.float64)

    def apply(self, sim):
        """Apply every block to the state of ``sim`` (a Simulator or its backend), one streaming sweep
        per block."""
        backend = getattr(sim, "backend", sim)
        for b in self.blocks:
            if max(b.qubits) >= backend.n_qubits:
                raise ValueError(f"block on qubits {b.qubits} does not fit a {backend.n_qubits}-qubit simulator")
        nq, qs, mats = self.arrays()
        _native.check(_native.lib().mqa_sim_apply_dense_blocks(backend._h, len(self.blocks), _native.ptr(nq),
                                                               _native.ptr(qs), _native.ptr(mats)))

    def as_circuit(self, keep_single_gates=True):
        """The blocks as a Circuit of UnivMathGate (blocks of at most 4 qubits) for the tile passes;
        a block of one gate stays that gate when ``keep_single_gates``."""
        circ = Circuit()
        for i, b in enumerate(self.blocks):
            if keep_single_gates and len(b.gates) == 1 and not b.gates[0].parameterized:
                circ += b.gates[0]
                continue
            if len(b.qubits) > 4:
                raise ValueError("the tile passes take dense blocks of at most 4 qubits")
            circ += G.UnivMathGate(f"fused{i}", b.matrix).on(list(b.qubits))
        return circ
