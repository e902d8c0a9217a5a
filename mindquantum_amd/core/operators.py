"""QubitOperator and Hamiltonian (Pauli-term form).

Host-side mirror of mindquantum/core/operators/qubit_operator.py and hamiltonian.py for the
``how_to == ORIGIN`` path: a Hamiltonian is a list of Pauli strings with real coefficients
(hamiltonian.py:90-101, 214-244).  ``terms_to_ir`` turns it into the (x, y, z, coeff) masks
that the kernels consume.
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


class Hamiltonian:
    """Pauli-term Hamiltonian with real coefficients (hamiltonian.py:35-244)."""

    def __init__(self, hamiltonian, dtype=None):
        from ..dtype import to_mq_type  # local import: dtype imports nothing from here

        if isinstance(hamiltonian, Hamiltonian):
            hamiltonian = hamiltonian.hamiltonian
        if not isinstance(hamiltonian, QubitOperator):
            raise TypeError(f"Hamiltonian requires a QubitOperator, but get {type(hamiltonian)}")
        for v in hamiltonian.terms.values():
            if abs(complex(v).imag) > 1e-12:
                raise ValueError("Hamiltonian requires real coefficients (a hermitian operator).")
        self.hamiltonian = hamiltonian
        self.dtype = to_mq_type(dtype if dtype is not None else "complex128")
        self.how_to = 0  # ORIGIN
        self.n_qubits = hamiltonian.count_qubits
        self._native = None

    def astype(self, dtype):
        return Hamiltonian(self.hamiltonian, dtype)

    def sparse(self, n_qubits=1):  # pragma: no cover - CSR Hamiltonians are not on this path
        raise NotImplementedError("sparse (CSR) Hamiltonians are not supported by the HIP backend yet")

    def terms_ir(self):
        """(x, y, z, coeff) masks of every term, as the C struct array of include/mqamd.h."""
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

            terms = self.terms_ir()
            h = ctypes.c_void_p()
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
        return repr(self.hamiltonian)
