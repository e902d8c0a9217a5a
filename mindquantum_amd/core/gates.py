"""Quantum gates of the state-vector path.

Host-side mirror of mindquantum/core/gates/basicgate.py for the gates the backend executes:
fixed gates (X, Y, Z, H, S, T, SX, I, SWAP, ISWAP, CNOT, UnivMathGate) and one-parameter
rotations (RX, RY, RZ, PhaseShift, GlobalPhase, Rxx, Ryy, Rzz).  Matrix conventions follow
the reference: the matrix of a multi-qubit gate is little-endian in ``obj_qubits`` order
(bit i of the matrix index <=> obj_qubits[i], cpu_vector_core_matrix_gate.cpp:85-149), and
parameterised gates take the hermitian by negating the parameter (core/gates/basic.py:510-514).
"""
import copy
import numbers

import numpy as np

from .parameterresolver import ParameterResolver


def _qubits(q):
    if q is None:
        return []
    if isinstance(q, (int, np.integer)):
        return [int(q)]
    return [int(i) for i in q]


class BasicGate:
    """Gate with target (``obj_qubits``) and control (``ctrl_qubits``) qubits."""

    _ir_kind = None
    n_qubits = 1
    parameterized = False

    def __init__(self, name, n_qubits=1):
        self.name = name
        self.n_qubits = n_qubits
        self.obj_qubits = []
        self.ctrl_qubits = []

    def on(self, obj_qubits, ctrl_qubits=None):
        """A copy of this gate acting on ``obj_qubits`` controlled by ``ctrl_qubits``."""
        new = copy.deepcopy(self)
        obj = _qubits(obj_qubits)
        ctrl = _qubits(ctrl_qubits)
        if len(obj) != self.n_qubits:
            raise ValueError(f"{self.name} gate requires {self.n_qubits} qubit(s), but get {len(obj)}.")
        if len(set(obj)) != len(obj):
            raise ValueError(f"Repeat qubits in obj qubits of {self.name}.")
        if set(obj) & set(ctrl):
            raise ValueError("Obj qubit and ctrl qubit cannot have same qubits.")
        if len(set(ctrl)) != len(ctrl):
            raise ValueError("Repeat qubits in ctrl qubits.")
        if any(q < 0 for q in obj + ctrl):
            raise ValueError("qubit index must be non-negative")
        new.obj_qubits = obj
        new.ctrl_qubits = ctrl
        return new

    def hermitian(self):
        return copy.deepcopy(self)

    @property
    def acted(self):
        return bool(self.obj_qubits)

    def __repr__(self):
        s = self.name
        if self.parameterized or getattr(self, "coeff", None) is not None and not isinstance(self, NoneParameterGate):
            s += f"({self.coeff.expression()})"
        if self.obj_qubits:
            s += "(" + " ".join(map(str, self.obj_qubits))
            if self.ctrl_qubits:
                s += " <-: " + " ".join(map(str, self.ctrl_qubits))
            s += ")"
        return s


class NoneParameterGate(BasicGate):
    """Gate with a fixed matrix."""

    _matrix = None

    def __init__(self, name, n_qubits=1, matrix=None, ir_kind=None, herm_kind=None):
        super().__init__(name, n_qubits)
        if matrix is not None:
            self._matrix = np.asarray(matrix, dtype=np.complex128)
        self._ir_kind = ir_kind
        self._herm_kind = herm_kind
        self.hermitianed = False

    def matrix(self, *args, **kwargs):  # pylint: disable=unused-argument
        m = self._matrix
        return m.conj().T if self.hermitianed else m.copy()

    def hermitian(self):
        new = copy.deepcopy(self)
        if self._herm_kind is not None:
            new.hermitianed = not self.hermitianed
            if self.name.endswith("dag"):
                new.name = self.name[:-3]
            else:
                new.name = self.name + "dag"
        return new

    @property
    def ir_kind(self):
        if self.hermitianed and self._herm_kind is not None:
            return self._herm_kind
        return self._ir_kind


_S2 = np.sqrt(0.5)


class XGate(NoneParameterGate):
    def __init__(self):
        super().__init__("X", 1, [[0, 1], [1, 0]], "X")


class YGate(NoneParameterGate):
    def __init__(self):
        super().__init__("Y", 1, [[0, -1j], [1j, 0]], "Y")


class ZGate(NoneParameterGate):
    def __init__(self):
        super().__init__("Z", 1, [[1, 0], [0, -1]], "Z")


class HGate(NoneParameterGate):
    def __init__(self):
        super().__init__("H", 1, [[_S2, _S2], [_S2, -_S2]], "H")


class IGate(NoneParameterGate):
    def __init__(self):
        super().__init__("I", 1, [[1, 0], [0, 1]], "I")


class SGate(NoneParameterGate):
    def __init__(self):
        super().__init__("S", 1, [[1, 0], [0, 1j]], "S", "Sdag")


class TGate(NoneParameterGate):
    def __init__(self):
        super().__init__("T", 1, [[1, 0], [0, (1 + 1j) * _S2]], "T", "Tdag")


class SXGate(NoneParameterGate):
    def __init__(self):
        super().__init__("SX", 1, [[0.5 + 0.5j, 0.5 - 0.5j], [0.5 - 0.5j, 0.5 + 0.5j]], "SX", "SXdag")


class SWAPGate(NoneParameterGate):
    def __init__(self):
        super().__init__("SWAP", 2, [[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], "SWAP")


class ISWAPGate(NoneParameterGate):
    def __init__(self):
        super().__init__("ISWAP", 2, [[1, 0, 0, 0], [0, 0, 1j, 0], [0, 1j, 0, 0], [0, 0, 0, 1]], "ISWAP", "ISWAPdag")


class CNOTGate(XGate):
    """CNOT: ``CNOT.on(obj, ctrl)`` or ``CNOT.on([obj, ctrl])`` (basicgate.py:401-456)."""

    def __init__(self):
        super().__init__()
        self.name = "CNOT"

    def on(self, obj_qubits, ctrl_qubits=None):
        obj = _qubits(obj_qubits)
        if ctrl_qubits is None:
            if len(obj) != 2:
                raise ValueError("CNOT.on needs [obj, ctrl] or obj and ctrl qubits")
            obj, ctrl_qubits = obj[:1], obj[1:]
        return XGate.on(self, obj, ctrl_qubits)


class UnivMathGate(NoneParameterGate):
    """Gate with a user matrix on up to 4 qubits (core/gates/basic.py UnivMathGate)."""

    def __init__(self, name, matrix_value):
        m = np.asarray(matrix_value, dtype=np.complex128)
        n = int(round(np.log2(m.shape[0])))
        if m.shape != (2**n, 2**n) or n < 1:
            raise ValueError("matrix_value must be a 2^n x 2^n matrix")
        if n > 4:
            raise ValueError("UnivMathGate supports at most 4 target qubits")
        super().__init__(name, n, m, "MATRIX")

    def hermitian(self):
        new = copy.deepcopy(self)
        new._matrix = self._matrix.conj().T.copy()
        new.name = self.name[:-3] if self.name.endswith("dag") else self.name + "dag"
        return new


class ParameterGate(BasicGate):
    """Gate whose matrix depends on one affine parameter ``coeff``."""

    def __init__(self, name, n_qubits, pr):
        super().__init__(name, n_qubits)
        if isinstance(pr, ParameterResolver):
            self.coeff = ParameterResolver(pr)
        elif isinstance(pr, (str, dict, numbers.Number)):
            self.coeff = ParameterResolver(pr)
        else:
            raise TypeError(f"Unsupported parameter type {type(pr)}")
        if self.coeff.is_complex:
            raise ValueError(f"{name} gate only supports real parameters")

    @property
    def parameterized(self):
        return not self.coeff.is_const()

    def _value(self, pr):
        if self.coeff.is_const():
            return float(np.real(self.coeff.const))
        if pr is None:
            raise ValueError("parameterized gate needs a parameter resolver")
        if isinstance(pr, numbers.Number):
            if len(self.coeff) != 1:
                raise ValueError("a scalar only resolves a one-parameter gate")
            return float(self.coeff.const + list(self.coeff.values())[0] * pr)
        return float(np.real(self.coeff.combination(pr)))

    def hermitian(self):
        new = copy.deepcopy(self)
        new.coeff = -self.coeff
        return new

    def matrix(self, pr=None):
        return self._mat(self._value(pr))

    def diff_matrix(self, pr=None):
        """Derivative of the matrix with respect to the angle value."""
        return self._dmat(self._value(pr))

    def no_grad(self):
        self.coeff.no_grad()
        return self

    def requires_grad(self):
        self.coeff.requires_grad()
        return self


def _rot(p_mat, th):
    """exp(-i th/2 P) = cos(th/2) I - i sin(th/2) P."""
    d = p_mat.shape[0]
    return np.cos(th / 2) * np.eye(d) - 1j * np.sin(th / 2) * p_mat


def _drot(p_mat, th):
    d = p_mat.shape[0]
    return -0.5 * np.sin(th / 2) * np.eye(d) - 0.5j * np.cos(th / 2) * p_mat


_PX = np.array([[0, 1], [1, 0]], dtype=np.complex128)
_PY = np.array([[0, -1j], [1j, 0]], dtype=np.complex128)
_PZ = np.array([[1, 0], [0, -1]], dtype=np.complex128)


class RotationGate(ParameterGate):
    _pauli = None
    _ir_kind = None

    def _mat(self, th):
        return _rot(self._pauli, th)

    def _dmat(self, th):
        return _drot(self._pauli, th)

    @property
    def ir_kind(self):
        return self._ir_kind


class RX(RotationGate):
    _pauli = _PX
    _ir_kind = "RX"

    def __init__(self, pr):
        super().__init__("RX", 1, pr)


class RY(RotationGate):
    _pauli = _PY
    _ir_kind = "RY"

    def __init__(self, pr):
        super().__init__("RY", 1, pr)


class RZ(RotationGate):
    _pauli = _PZ
    _ir_kind = "RZ"

    def __init__(self, pr):
        super().__init__("RZ", 1, pr)


class Rxx(RotationGate):
    _pauli = np.kron(_PX, _PX)
    _ir_kind = "Rxx"

    def __init__(self, pr):
        super().__init__("Rxx", 2, pr)


class Ryy(RotationGate):
    _pauli = np.kron(_PY, _PY)
    _ir_kind = "Ryy"

    def __init__(self, pr):
        super().__init__("Ryy", 2, pr)


class Rzz(RotationGate):
    _pauli = np.kron(_PZ, _PZ)
    _ir_kind = "Rzz"

    def __init__(self, pr):
        super().__init__("Rzz", 2, pr)


class PhaseShift(ParameterGate):
    """diag(1, e^{i th}) (basicgate.py:1676-1716)."""

    ir_kind = "PS"

    def __init__(self, pr):
        super().__init__("PS", 1, pr)

    def _mat(self, th):
        return np.array([[1, 0], [0, np.exp(1j * th)]], dtype=np.complex128)

    def _dmat(self, th):
        return np.array([[0, 0], [0, 1j * np.exp(1j * th)]], dtype=np.complex128)


class GlobalPhase(ParameterGate):
    """exp(-i th) I (basicgate.py:1495-1535)."""

    ir_kind = "GP"

    def __init__(self, pr):
        super().__init__("GP", 1, pr)

    def _mat(self, th):
        return np.exp(-1j * th) * np.eye(2, dtype=np.complex128)

    def _dmat(self, th):
        return -1j * np.exp(-1j * th) * np.eye(2, dtype=np.complex128)


X = XGate()
Y = YGate()
Z = ZGate()
H = HGate()
S = SGate()
T = TGate()
SX = SXGate()
I = IGate()  # noqa: E741
SWAP = SWAPGate()
ISWAP = ISWAPGate()
CNOT = CNOTGate()

__all__ = [
    "BasicGate", "NoneParameterGate", "ParameterGate", "RotationGate",
    "X", "Y", "Z", "H", "S", "T", "SX", "I", "SWAP", "ISWAP", "CNOT",
    "XGate", "YGate", "ZGate", "HGate", "SGate", "TGate", "SXGate", "IGate", "SWAPGate", "ISWAPGate", "CNOTGate",
    "UnivMathGate", "RX", "RY", "RZ", "Rxx", "Ryy", "Rzz", "PhaseShift", "GlobalPhase",
]
