"""QubitOperator and Hamiltonian (Pauli-term form).

Host-side mirror of mindquantum/core/operators/qubit_operator.py and hamiltonian.py: a Hamiltonian
is a list of Pauli strings with real coefficients (hamiltonian.py:90-101, 214-244; ``terms_ir``
turns it into the (x, y, z, coeff) masks that the kernels consume), or a CSR sparse matrix
(``how_to == FRONTEND``).
"""
import numbers
import re

import numpy as np

from .. import _native

_TERM_RE = re.compile(r"([XYZxyz])(\d+)")


def _parse(term):
    if term is None or term == "":
        return 1 + 0j, ()
    if isinstance(term, str):
        out = []
        for tok in term.split():
            m = _TERM_RE.fullmatch(tok)
            if not m:
                raise ValueError(f"Invalid pauli term {tok}")
            out.append((int(m.group(2)), m.group(1).upper()))
        return _normalise(out)
    if isinstance(term, tuple):
        return _normalise(list(term))
    raise TypeError(f"Unsupported term type {type(term)}")


def _normalise(ops):
    """Product of single-qubit Paulis sorted by qubit: returns (phase, ((q, P), ...))."""
    phase = 1 + 0j
    by_q = {}
    for q, p in ops:
        if q in by_q:
            ph, res = _mul1(by_q[q], p)
            phase *= ph
            if res == "I":
                del by_q[q]
            else:
                by_q[q] = res
        else:
            by_q[q] = p
    return phase, tuple(sorted(by_q.items()))


_MUL = {
    ("X", "Y"): (1j, "Z"), ("Y", "X"): (-1j, "Z"),
    ("Y", "Z"): (1j, "X"), ("Z", "Y"): (-1j, "X"),
    ("Z", "X"): (1j, "Y"), ("X", "Z"): (-1j, "Y"),
}


def _mul1(a, b):
    if a == b:
        return 1, "I"
    return _MUL[(a, b)]


class QubitOperator:
    """Sum of Pauli strings with complex coefficients, e.g. ``QubitOperator('X0 Y1', 0.5)``."""

    def __init__(self, term=None, coefficient=1.0):
        self.terms = {}
        if term is None:
            return
        phase, key = _parse(term)
        self.terms[key] = phase * coefficient

    @classmethod
    def _from(cls, terms):
        out = cls()
        out.terms = {k: v for k, v in terms.items() if v != 0}
        return out

    def __add__(self, other):
        if isinstance(other, numbers.Number):
            other = QubitOperator("", other)
        out = dict(self.terms)
        for k, v in other.terms.items():
            out[k] = out.get(k, 0) + v
        return QubitOperator._from(out)

    __radd__ = __add__

    def __neg__(self):
        return QubitOperator._from({k: -v for k, v in self.terms.items()})

    def __sub__(self, other):
        return self + (-other)

    def __rsub__(self, other):
        return (-self) + other

    def __mul__(self, other):
        if isinstance(other, numbers.Number):
            return QubitOperator._from({k: v * other for k, v in self.terms.items()})
        out = {}
        for k1, v1 in self.terms.items():
            for k2, v2 in other.terms.items():
                phase, key = _normalise(list(k1) + list(k2))
                out[key] = out.get(key, 0) + v1 * v2 * phase
        return QubitOperator._from(out)

    def __rmul__(self, other):
        return self * other

    def __truediv__(self, other):
        return self * (1.0 / other)

    @property
    def real(self):
        return QubitOperator._from({k: complex(v).real for k, v in self.terms.items()})

    @property
    def imag(self):
        return QubitOperator._from({k: complex(v).imag for k, v in self.terms.items()})

    @property
    def count_qubits(self):
        qs = [q for k in self.terms for q, _ in k]
        return max(qs) + 1 if qs else 0

    def __len__(self):
        return len(self.terms)

    def __repr__(self):
        parts = []
        for k, v in self.terms.items():
            s = " ".join(f"{p}{q}" for q, p in k)
            parts.append(f"{v} [{s}]")
        return " +\n".join(parts) if parts else "0"


ORIGIN, BACKEND, FRONTEND = 0, 1, 2  # HowTo of hamiltonian.py


class Hamiltonian:
    """Hamiltonian (hamiltonian.py:35-244): Pauli terms with real coefficients, or a
    ``scipy.sparse.csr_matrix`` of size 2^n x 2^n (``how_to == FRONTEND``: applied as a sparse
    matrix-vector product on the device).  ``sparse(n)`` on a Pauli Hamiltonian marks it
    ``BACKEND`` as the reference does; it keeps running on the Pauli-term engine, which computes the
    same operator."""

    def __init__(self, hamiltonian, dtype=None):
        import scipy.sparse as sp

        from ..dtype import to_mq_type  # local import: dtype imports nothing from here

        self._sparse = None
        if isinstance(hamiltonian, Hamiltonian):
            hamiltonian = hamiltonian._sparse if hamiltonian.how_to == FRONTEND else hamiltonian.hamiltonian
        if sp.issparse(hamiltonian):
            if not isinstance(hamiltonian, (sp.csr_matrix, sp.csr_array)):
                raise TypeError(f"Require a QubitOperator or a csr_matrix, but get {type(hamiltonian)}!")
            if len(hamiltonian.shape) != 2 or hamiltonian.shape[0] != hamiltonian.shape[1]:
                raise ValueError(f"Hamiltonian requires a two dimension square csr_matrix, but get shape "
                                 f"{hamiltonian.shape}")
            dim = hamiltonian.shape[0]
            if dim < 1 or dim & (dim - 1):
                raise ValueError(f"size of hamiltonian sparse matrix should be power of 2, but get {hamiltonian.shape}")
            self._sparse = sp.csr_matrix(hamiltonian, dtype=np.complex128)
            self._sparse.sort_indices()
            self.hamiltonian = None
            self.how_to = FRONTEND
            self.n_qubits = int(dim).bit_length() - 1
        else:
            if not isinstance(hamiltonian, QubitOperator):
                raise TypeError(f"Hamiltonian requires a QubitOperator, but get {type(hamiltonian)}")
            for v in hamiltonian.terms.values():
                if abs(complex(v).imag) > 1e-12:
                    raise ValueError("Hamiltonian requires real coefficients (a hermitian operator).")
            self.hamiltonian = hamiltonian
            self.how_to = ORIGIN
            self.n_qubits = hamiltonian.count_qubits
        self.dtype = to_mq_type(dtype if dtype is not None else "complex128")
        self._native = None

    def astype(self, dtype):
        out = Hamiltonian(self._sparse if self.how_to == FRONTEND else self.hamiltonian, dtype)
        if self.how_to == BACKEND:
            out.how_to, out.n_qubits = BACKEND, self.n_qubits
        return out

    def sparse(self, n_qubits=1):
        """Mark the Pauli Hamiltonian for the sparse backend form (hamiltonian.py:137-151)."""
        if self.how_to != ORIGIN:
            raise ValueError("Already a sparse hamiltonian.")
        if n_qubits < self.n_qubits:
            raise ValueError(f"Can not sparse a {self.n_qubits} qubits hamiltonian to {n_qubits} hamiltonian.")
        self.n_qubits = n_qubits
        self.how_to = BACKEND
        return self

    def sparse_matrix(self):
        """The Hamiltonian as a csr_matrix (little-endian qubit order)."""
        import scipy.sparse as sp

        if self.how_to == FRONTEND:
            return self._sparse.copy()
        n = max(self.n_qubits, 1)
        pauli = {"X": sp.csr_matrix([[0, 1], [1, 0]], dtype=complex), "Y": sp.csr_matrix([[0, -1j], [1j, 0]]),
                 "Z": sp.csr_matrix([[1, 0], [0, -1]], dtype=complex)}
        out = sp.csr_matrix((1 << n, 1 << n), dtype=np.complex128)
        for key, coeff in self.hamiltonian.terms.items():
            ops = dict((q, p) for q, p in key)
            m = sp.identity(1, dtype=np.complex128, format="csr")
            for q in range(n - 1, -1, -1):
                m = sp.kron(m, pauli[ops[q]] if q in ops else sp.identity(2, format="csr"), format="csr")
            out = out + complex(coeff) * m
        return out.tocsr()

    def terms_ir(self):
        """(x, y, z, coeff) masks of every term, as the C struct array of include/mqamd.h."""
        if self.how_to == FRONTEND:
            raise ValueError("this Hamiltonian was constructed from a sparse matrix; it has no Pauli terms")
        arr = np.zeros(len(self.hamiltonian.terms), dtype=_native.TERM_DTYPE)
        for i, (k, v) in enumerate(self.hamiltonian.terms.items()):
            for q, p in k:
                field = {"X": "x_mask", "Y": "y_mask", "Z": "z_mask"}[p]
                arr[i][field] = int(arr[i][field]) | (1 << q)
            arr[i]["coeff"] = float(complex(v).real)
        return arr

    def native(self):
        """Handle of the C ABI Hamiltonian (created once)."""
        if self._native is None:
            import ctypes

            h = ctypes.c_void_p()
            if self.how_to == FRONTEND:
                m = self._sparse
                ptr = np.ascontiguousarray(m.indptr, dtype=np.int64)
                idx = np.ascontiguousarray(m.indices, dtype=np.int64)
                val = np.ascontiguousarray(m.data, dtype=np.complex128).view(np.float64)
                _native.check(_native.lib().mqa_ham_create_csr(m.shape[0], m.nnz, _native.ptr(ptr),
                                                                _native.ptr(idx) if m.nnz else None,
                                                                _native.ptr(val) if m.nnz else None,
                                                                ctypes.byref(h)))
            else:
                terms = self.terms_ir()
                _native.check(_native.lib().mqa_ham_create(_native.ptr(terms), len(terms), ctypes.byref(h)))
            self._native = h
        return self._native

    def __del__(self):
        h = getattr(self, "_native", None)
        if h is not None and _native._lib is not None:
            _native._lib.mqa_ham_destroy(h)
            self._native = None

    def __getstate__(self):
        d = dict(self.__dict__)
        d["_native"] = None
        return d

    def __repr__(self):
        return repr(self._sparse) if self.how_to == FRONTEND else repr(self.hamiltonian)
