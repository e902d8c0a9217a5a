"""Quantum gates of the state-vector path.

Host-side mirror of mindquantum/core/gates/basicgate.py for the gates the backend executes:
fixed gates (X, Y, Z, H, S, T, SX, I, SWAP, ISWAP, CNOT, UnivMathGate, GroupedPauli, Power,
Barrier), one-parameter rotations (RX, RY, RZ, PhaseShift, GlobalPhase, Rxx, Ryy, Rzz, Rxy, Rxz,
Ryz, RotPauliString, Givens, SWAPalpha) and multi-parameter gates (U3, FSim, constant Rn).  Gates
the kernels do not run directly are lowered (``lowered()``) into an exact product of gates they do
run, with the same controls and affine parameter expressions, so derivative taps stay exact.  Matrix conventions follow
the reference: the matrix of a multi-qubit gate is little-endian in ``obj_qubits`` order
(bit i of the matrix index <=> obj_qubits[i], cpu_vector_core_matrix_gate.cpp:85-149), and
parameterised gates take the hermitian by negating the parameter (core/gates/basic.py:510-514).
"""
import copy
import numbers

import numpy as np

from .mutation import bump
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

    def __setattr__(self, key, value):
        # any change of a gate can change the native circuits lowered from circuits that hold it
        object.__setattr__(self, key, value)
        bump()

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

    def param_resolvers(self):
        """The gate's parameter expressions (one per parameter slot of its matrix)."""
        return []

    def lowered(self):
        """The gate as a sequence of gates the backend executes directly (the gate itself for
        those); composite gates return an exact decomposition with the same controls."""
        return [self]

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

    def param_resolvers(self):
        return [self.coeff]

    def diff_matrices(self, pr=None):
        """Derivative of the matrix with respect to each parameter expression's value."""
        return [self.diff_matrix(pr)]

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


_PAULI_OF = {"X": _PX, "Y": _PY, "Z": _PZ, "I": np.eye(2, dtype=np.complex128)}


def _tag(gates, k, factor):
    """Record on lowered gates that their angle is ``factor`` times parameter expression k of the
    gate they came from (nested lowerings multiply factors); untagged parameterised gates carry
    constant angles.  Used to shift one expression of a composite gate (parameter-shift rule)."""
    out = []
    for g in gates:
        if isinstance(g, ParameterGate) and not g.coeff.is_const():
            inner = getattr(g, "_origin", (0, 1.0))
            g._origin = (k, factor * inner[1])
        out.append(g)
    return out


def _kron_string(paulis):
    """Matrix of a Pauli string, paulis[i] acting on matrix bit i (little-endian)."""
    m = np.ones((1, 1), dtype=np.complex128)
    for p in reversed(paulis):
        m = np.kron(m, _PAULI_OF[p])
    return m


def _basis_to_z(p, q, inverse):
    """Fixed gate that maps Pauli p on qubit q to Z (inverse: back), or None for Z.
    H X H = Z; RX(pi/2) Y RX(-pi/2) = Z."""
    if p == "X":
        return H.on(q)
    if p == "Y":
        return RX(-np.pi / 2 if inverse else np.pi / 2).on(q)
    return None


def _pauli_rotation(paulis, objs, ctrl, pr):
    """exp(-i pr/2 P) for the Pauli string P = paulis[i] on objs[i] (identities dropped): basis
    change to Z...Z, then RZ / Rzz on the parity (CNOT ladder for more than two qubits).  Only the
    rotation carries the controls: V^dag C(R) V is C(V^dag R V) for any V."""
    act = [(p, q) for p, q in zip(paulis, objs) if p != "I"]
    if not act:
        return _tag([GlobalPhase(pr * 0.5).on(objs[0], ctrl)], 0, 0.5)
    if len(act) == 1:
        p, q = act[0]
        return _tag([{"X": RX, "Y": RY, "Z": RZ}[p](pr).on(q, ctrl)], 0, 1.0)
    pre = [g for g in (_basis_to_z(p, q, False) for p, q in act) if g is not None]
    post = [g for g in (_basis_to_z(p, q, True) for p, q in act) if g is not None]
    qs = [q for _, q in act]
    if len(qs) == 2:
        core = _tag([Rzz(pr).on(qs, ctrl)], 0, 1.0)
    else:
        ladder = [X.on(qs[i + 1], qs[i]) for i in range(len(qs) - 2)]
        core = ladder + _tag([Rzz(pr).on(qs[-2:], ctrl)], 0, 1.0) + ladder[::-1]
    return pre + core + post


class PauliRotationGate(ParameterGate):
    """exp(-i th/2 P) for a fixed Pauli string P (``_paulis[i]`` on obj_qubits[i])."""

    _paulis = ""

    def _mat(self, th):
        return _rot(_kron_string(self._paulis), th)

    def _dmat(self, th):
        return _drot(_kron_string(self._paulis), th)

    def lowered(self):
        out = []
        for g in _pauli_rotation(self._paulis, self.obj_qubits, self.ctrl_qubits, self.coeff):
            out.extend(g.lowered())
        return out


class Rxy(PauliRotationGate):
    """exp(-i th/2 X_{q0} Y_{q1}) (basicgate.py:1201-1281, core PauliStringGate([X, Y]))."""

    _paulis = "XY"

    def __init__(self, pr):
        super().__init__("Rxy", 2, pr)


class Rxz(PauliRotationGate):
    """exp(-i th/2 X_{q0} Z_{q1}) (basicgate.py:1283-1361)."""

    _paulis = "XZ"

    def __init__(self, pr):
        super().__init__("Rxz", 2, pr)


class Ryz(PauliRotationGate):
    """exp(-i th/2 Y_{q0} Z_{q1}) (basicgate.py:1363-1441)."""

    _paulis = "YZ"

    def __init__(self, pr):
        super().__init__("Ryz", 2, pr)


class RotPauliString(PauliRotationGate):
    """exp(-i th/2 P) for an arbitrary Pauli string, character i on obj_qubits[i]
    (basicgate.py:1567-1674)."""

    def __init__(self, pauli_string, pr):
        if not isinstance(pauli_string, str) or any(c.upper() not in "IXYZ" for c in pauli_string):
            raise ValueError(f"Unknown pauli string {pauli_string!r}")
        self._paulis = pauli_string.upper()
        super().__init__("RPS", len(pauli_string), pr)


class Givens(ParameterGate):
    """[[1,0,0,0],[0,cos,-sin,0],[0,sin,cos,0],[0,0,0,1]] = exp(-i th/2 (Y X - X Y))
    (basicgate.py:617-704).  Y X (kron) is X on q0, Y on q1; the two terms commute, so the gate
    is Rxy(th) on (q0, q1) followed by Rxy(-th) on (q1, q0)."""

    def __init__(self, pr):
        super().__init__("Givens", 2, pr)

    def _mat(self, th):
        c, s = np.cos(th), np.sin(th)
        return np.array([[1, 0, 0, 0], [0, c, -s, 0], [0, s, c, 0], [0, 0, 0, 1]], dtype=np.complex128)

    def _dmat(self, th):
        c, s = np.cos(th), np.sin(th)
        return np.array([[0, 0, 0, 0], [0, -s, -c, 0], [0, c, -s, 0], [0, 0, 0, 0]], dtype=np.complex128)

    def lowered(self):
        a, b = self.obj_qubits
        return _tag(Rxy(self.coeff).on([a, b], self.ctrl_qubits).lowered(), 0, 1.0) + \
            _tag(Rxy(-self.coeff).on([b, a], self.ctrl_qubits).lowered(), 0, -1.0)


class SWAPalpha(ParameterGate):
    """SWAP^alpha: e = exp(i pi alpha), [[1,0,0,0],[0,(1+e)/2,(1-e)/2,0],[0,(1-e)/2,(1+e)/2,0],
    [0,0,0,1]] (basicgate.py:849-937).  With SWAP = (I + XX + YY + ZZ)/2 it equals
    exp(i pi alpha/4) exp(-i pi alpha/4 XX) exp(-i pi alpha/4 YY) exp(-i pi alpha/4 ZZ), i.e.
    GlobalPhase(-pi alpha/4) Rxx(pi alpha/2) Ryy(pi alpha/2) Rzz(pi alpha/2)."""

    def __init__(self, pr):
        super().__init__("SWAPalpha", 2, pr)

    def _mat(self, a):
        e = np.exp(1j * np.pi * a)
        return np.array([[1, 0, 0, 0], [0, (1 + e) / 2, (1 - e) / 2, 0], [0, (1 - e) / 2, (1 + e) / 2, 0],
                         [0, 0, 0, 1]], dtype=np.complex128)

    def _dmat(self, a):
        e = 0.5j * np.pi * np.exp(1j * np.pi * a)
        return np.array([[0, 0, 0, 0], [0, e, -e, 0], [0, -e, e, 0], [0, 0, 0, 0]], dtype=np.complex128)

    def lowered(self):
        half = self.coeff * (np.pi / 2)
        q, c = self.obj_qubits, self.ctrl_qubits
        return _tag([GlobalPhase(self.coeff * (-np.pi / 4)).on(q[0], c)], 0, -np.pi / 4) + \
            _tag([Rxx(half).on(q, c), Ryy(half).on(q, c), Rzz(half).on(q, c)], 0, np.pi / 2)


class MultiParamsGate(BasicGate):
    """Gate whose matrix depends on several affine parameter expressions ``prs``
    (basicgate.py:1962-2230)."""

    def __init__(self, name, n_qubits, prs):
        super().__init__(name, n_qubits)
        self.prs = []
        for pr in prs:
            if not isinstance(pr, (ParameterResolver, str, dict, numbers.Number)):
                raise TypeError(f"Unsupported parameter type {type(pr)}")
            p = ParameterResolver(pr)
            if p.is_complex:
                raise ValueError(f"{name} gate only supports real parameters")
            self.prs.append(p)

    @property
    def parameterized(self):
        return any(not p.is_const() for p in self.prs)

    def param_resolvers(self):
        return self.prs

    def _values(self, pr):
        out = []
        for p in self.prs:
            if p.is_const():
                out.append(float(np.real(p.const)))
            elif pr is None:
                raise ValueError("parameterized gate needs a parameter resolver")
            else:
                out.append(float(np.real(p.combination(pr))))
        return out

    def matrix(self, pr=None):
        return self._mat(*self._values(pr))

    def diff_matrices(self, pr=None):
        return self._dmats(*self._values(pr))

    def no_grad(self):
        for p in self.prs:
            p.no_grad()
        return self

    def requires_grad(self):
        for p in self.prs:
            p.requires_grad()
        return self

    def __repr__(self):
        s = self.name + "(" + ", ".join(p.expression() for p in self.prs) + ")"
        if self.obj_qubits:
            s += "(" + " ".join(map(str, self.obj_qubits))
            if self.ctrl_qubits:
                s += " <-: " + " ".join(map(str, self.ctrl_qubits))
            s += ")"
        return s


class U3(MultiParamsGate):
    """[[cos(t/2), -e^{i l} sin(t/2)], [e^{i p} sin(t/2), e^{i(p+l)} cos(t/2)]]
    (basicgate.py:2370-2510) = PhaseShift(p) RY(t) PhaseShift(l)."""

    def __init__(self, theta, phi, lamda):
        super().__init__("U3", 1, [theta, phi, lamda])

    @property
    def theta(self):
        return self.prs[0]

    @property
    def phi(self):
        return self.prs[1]

    @property
    def lamda(self):
        return self.prs[2]

    def hermitian(self):
        """U3(t, p, l)^dag = U3(-t, -l, -p) (basicgate.py:2445-2458)."""
        out = U3(-self.prs[0], -self.prs[2], -self.prs[1])
        out.obj_qubits, out.ctrl_qubits = list(self.obj_qubits), list(self.ctrl_qubits)
        return out

    def _mat(self, t, p, l):
        c, s = np.cos(t / 2), np.sin(t / 2)
        return np.array([[c, -np.exp(1j * l) * s], [np.exp(1j * p) * s, np.exp(1j * (p + l)) * c]],
                        dtype=np.complex128)

    def _dmats(self, t, p, l):
        c, s = np.cos(t / 2), np.sin(t / 2)
        el, ep, epl = np.exp(1j * l), np.exp(1j * p), np.exp(1j * (p + l))
        d_t = np.array([[-s / 2, -el * c / 2], [ep * c / 2, -epl * s / 2]], dtype=np.complex128)
        d_p = np.array([[0, 0], [1j * ep * s, 1j * epl * c]], dtype=np.complex128)
        d_l = np.array([[0, -1j * el * s], [0, 1j * epl * c]], dtype=np.complex128)
        return [d_t, d_p, d_l]

    def lowered(self):
        q, c = self.obj_qubits, self.ctrl_qubits
        return _tag([PhaseShift(self.prs[2]).on(q, c)], 2, 1.0) + _tag([RY(self.prs[0]).on(q, c)], 0, 1.0) + \
            _tag([PhaseShift(self.prs[1]).on(q, c)], 1, 1.0)


class FSim(MultiParamsGate):
    """[[1,0,0,0],[0,cos t,-i sin t,0],[0,-i sin t,cos t,0],[0,0,0,e^{-i p}]]
    (basicgate.py:2512-2600).  (XX + YY)/2 swaps |01> and |10> and annihilates |00>, |11>, so the
    middle block is Rxx(t) Ryy(t); the phase on |11> is PhaseShift(-p) on q1 controlled by q0."""

    def __init__(self, theta, phi):
        super().__init__("FSim", 2, [theta, phi])

    @property
    def theta(self):
        return self.prs[0]

    @property
    def phi(self):
        return self.prs[1]

    def hermitian(self):
        out = FSim(-self.prs[0], -self.prs[1])
        out.obj_qubits, out.ctrl_qubits = list(self.obj_qubits), list(self.ctrl_qubits)
        return out

    def _mat(self, t, p):
        a, b = np.cos(t), -1j * np.sin(t)
        return np.array([[1, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, np.exp(-1j * p)]], dtype=np.complex128)

    def _dmats(self, t, p):
        a, b = -np.sin(t), -1j * np.cos(t)
        d_t = np.array([[0, 0, 0, 0], [0, a, b, 0], [0, b, a, 0], [0, 0, 0, 0]], dtype=np.complex128)
        d_p = np.zeros((4, 4), dtype=np.complex128)
        d_p[3, 3] = -1j * np.exp(-1j * p)
        return [d_t, d_p]

    def lowered(self):
        q, c = self.obj_qubits, list(self.ctrl_qubits)
        return _tag([Rxx(self.prs[0]).on(q, c), Ryy(self.prs[0]).on(q, c)], 0, 1.0) + \
            _tag([PhaseShift(-self.prs[1]).on(q[1], [q[0]] + c)], 1, -1.0)


class Rn(MultiParamsGate):
    """exp(-i (a X + b Y + c Z)/2) (basicgate.py:2232-2368).  The rotation axis depends on the
    parameters non-linearly, so only constant Rn gates run here (as a dense matrix)."""

    def __init__(self, alpha, beta, gamma):
        super().__init__("Rn", 1, [alpha, beta, gamma])

    def hermitian(self):
        out = Rn(-self.prs[0], -self.prs[1], -self.prs[2])
        out.obj_qubits, out.ctrl_qubits = list(self.obj_qubits), list(self.ctrl_qubits)
        return out

    def _mat(self, a, b, c):
        f = np.sqrt(a * a + b * b + c * c)
        if f == 0:
            return np.eye(2, dtype=np.complex128)
        gen = (a * _PX + b * _PY + c * _PZ) / f
        return np.cos(f / 2) * np.eye(2) - 1j * np.sin(f / 2) * gen

    def _dmats(self, a, b, c):
        """d/dv_j of cos(f/2) I - i sin(f/2) n.sigma with v = (a, b, c), f = |v|, n = v / f:
        -(sin(f/2) / 2) n_j I - i [(cos(f/2) / 2) n_j n.sigma + sin(f/2) (sigma_j - n_j n.sigma) / f];
        at f = 0: -i sigma_j / 2."""
        v = np.array([a, b, c], dtype=float)
        f = float(np.sqrt(v @ v))
        sig = (_PX, _PY, _PZ)
        if f == 0:
            return [-0.5j * p for p in sig]
        nv = v / f
        ns = nv[0] * _PX + nv[1] * _PY + nv[2] * _PZ
        cf, sf = np.cos(f / 2), np.sin(f / 2)
        return [-0.5 * sf * nv[j] * np.eye(2) - 1j * (0.5 * cf * nv[j] * ns + sf * (sig[j] - nv[j] * ns) / f)
                for j in range(3)]

    @property
    def ir_kind(self):
        return "PMATRIX"

    def lowered(self):
        if self.parameterized:  # per-sample matrices from the closed form above
            return [self]
        m = UnivMathGate("Rn", self.matrix())
        return [m.on(self.obj_qubits, self.ctrl_qubits)]


class ParamCustomGate(ParameterGate):
    """A gate of one parameter whose matrix is a user function (basicgate.py:1828-1960
    gene_univ_parameterized_gate): ``matrix_generator(x)`` and ``diff_matrix_generator(x)`` (its
    derivative) return 2^k x 2^k complex arrays.  The hermitian gate keeps the parameter and takes the
    conjugate transposes (basicgate.py:1905-1911, 1937-1945).  The backend evaluates the functions on
    the host per sample and runs the matrices as dense ops (k <= 2)."""

    _gen = None
    _dgen = None

    def __init__(self, pr):
        super().__init__(self._name, self._n_qubits, pr)
        self.hermitianed = False

    def _mat(self, x):
        m = np.asarray(self._gen(x), dtype=np.complex128)
        return m.conj().T.copy() if self.hermitianed else m

    def _dmat(self, x):
        m = np.asarray(self._dgen(x), dtype=np.complex128)
        return m.conj().T.copy() if self.hermitianed else m

    def hermitian(self):
        new = copy.deepcopy(self)
        new.hermitianed = not self.hermitianed
        return new

    @property
    def ir_kind(self):
        return "PMATRIX"


def _check_generator(fn, n_args, what):
    import inspect

    if len(inspect.signature(fn).parameters) != n_args:
        raise ValueError(f"{what} can only have {n_args} argument{'s' if n_args > 1 else ''}, but get "
                         f"{len(inspect.signature(fn).parameters)}.")


def gene_univ_parameterized_gate(name, matrix_generator, diff_matrix_generator):
    """A gate class of one parameter with a user matrix and derivative (basicgate.py:1828-1960): the
    same checks on the two generators, no numba needed."""
    _check_generator(matrix_generator, 1, "matrix_generator")
    _check_generator(diff_matrix_generator, 1, "diff_matrix_generator")
    m, dm = matrix_generator(0.123), diff_matrix_generator(0.123)
    if not isinstance(m, np.ndarray) or m.dtype != complex:
        raise ValueError(f"matrix_generator should return numpy array with complex type, but get {type(m)}.")
    if not isinstance(dm, np.ndarray) or dm.dtype != complex:
        raise ValueError(f"diff_matrix_generator should return numpy array with complex type, but get {type(dm)}")
    if m.shape != dm.shape:
        raise ValueError("matrix_generator and diff_matrix_generator should generate same shape matrix.")
    n = int(round(np.log2(m.shape[0])))
    return type(name, (ParamCustomGate,), {"_name": name, "_n_qubits": n, "_gen": staticmethod(matrix_generator),
                                           "_dgen": staticmethod(diff_matrix_generator)})


class TwoParamCustomGate(BasicGate):
    """A gate of two parameters with user matrix functions (basicgate.py:2039-2230
    gene_univ_two_params_gate): ``matrix_generator(a, b)`` and its two partial derivatives."""

    _gen = None
    _dgens = ()

    def __init__(self, pr1, pr2):
        super().__init__(self._name, self._n_qubits)
        self.prs = []
        for pr in (pr1, pr2):
            p = ParameterResolver(pr)
            if p.is_complex:
                raise ValueError(f"{self._name} gate only supports real parameters")
            self.prs.append(p)
        self.hermitianed = False

    @property
    def parameterized(self):
        return any(not p.is_const() for p in self.prs)

    def param_resolvers(self):
        return self.prs

    def _values(self, pr):
        return MultiParamsGate._values(self, pr)

    def matrix(self, pr=None):
        m = np.asarray(self._gen(*self._values(pr)), dtype=np.complex128)
        return m.conj().T.copy() if self.hermitianed else m

    def diff_matrices(self, pr=None):
        vals = self._values(pr)
        out = [np.asarray(d(*vals), dtype=np.complex128) for d in self._dgens]
        return [d.conj().T.copy() for d in out] if self.hermitianed else out

    def hermitian(self):
        new = copy.deepcopy(self)
        new.hermitianed = not self.hermitianed
        return new

    def no_grad(self):
        for p in self.prs:
            p.no_grad()
        return self

    def requires_grad(self):
        for p in self.prs:
            p.requires_grad()
        return self

    @property
    def ir_kind(self):
        return "PMATRIX"


def gene_univ_two_params_gate(name, matrix_generator, diff_matrix_generator_1, diff_matrix_generator_2):
    """A gate class of two parameters with a user matrix and its partial derivatives
    (basicgate.py:2039-2230), no numba needed."""
    _check_generator(matrix_generator, 2, "matrix_generator")
    _check_generator(diff_matrix_generator_1, 2, "diff_matrix_generator_1")
    _check_generator(diff_matrix_generator_2, 2, "diff_matrix_generator_2")
    m = matrix_generator(0.123, 0.456)
    for d in (diff_matrix_generator_1(0.123, 0.456), diff_matrix_generator_2(0.123, 0.456)):
        if not isinstance(d, np.ndarray) or d.shape != np.shape(m):
            raise ValueError("matrix_generator and the diff matrix generators should generate same shape matrix.")
    n = int(round(np.log2(np.shape(m)[0])))
    return type(name, (TwoParamCustomGate,), {
        "_name": name, "_n_qubits": n, "_gen": staticmethod(matrix_generator),
        "_dgens": (diff_matrix_generator_1, diff_matrix_generator_2)})


class GroupedPauli(NoneParameterGate):
    """Tensor product of Paulis, character i on obj_qubits[i] (basicgate.py:126-202)."""

    def __init__(self, pauli_string):
        if not isinstance(pauli_string, str) or any(c.upper() not in "IXYZ" for c in pauli_string):
            raise ValueError(f"Unknown pauli string {pauli_string!r}")
        self.pauli_string = pauli_string.upper()
        super().__init__("GroupedPauli", len(pauli_string), _kron_string(self.pauli_string))

    def lowered(self):
        return [{"X": X, "Y": Y, "Z": Z}[p].on(q, self.ctrl_qubits)
                for p, q in zip(self.pauli_string, self.obj_qubits) if p != "I"]


class Power(UnivMathGate):
    """A fixed gate's matrix raised to ``exponent`` (basicgate.py:1756-1800)."""

    def __init__(self, gate, exponent=0.5):
        from scipy.linalg import fractional_matrix_power

        if not isinstance(gate, NoneParameterGate):
            raise TypeError("Power needs a non-parameterized gate")
        super().__init__(f"{gate.name}^{exponent}", fractional_matrix_power(gate.matrix(), exponent))
        self.gate = gate
        self.t = exponent


class BarrierGate(BasicGate):
    """No operation (basicgate.py:1443-1493)."""

    def __init__(self, show=True):
        super().__init__("BARRIER", 0)
        self.show = show

    def on(self, obj_qubits, ctrl_qubits=None):
        new = copy.deepcopy(self)
        new.obj_qubits = _qubits(obj_qubits)
        return new

    def lowered(self):
        return []


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
FSIM = FSim  # reference alias

__all__ = [
    "BasicGate", "NoneParameterGate", "ParameterGate", "RotationGate",
    "X", "Y", "Z", "H", "S", "T", "SX", "I", "SWAP", "ISWAP", "CNOT",
    "XGate", "YGate", "ZGate", "HGate", "SGate", "TGate", "SXGate", "IGate", "SWAPGate", "ISWAPGate", "CNOTGate",
    "UnivMathGate", "RX", "RY", "RZ", "Rxx", "Ryy", "Rzz", "PhaseShift", "GlobalPhase",
    "PauliRotationGate", "Rxy", "Rxz", "Ryz", "RotPauliString", "Givens", "SWAPalpha", "MultiParamsGate", "U3", "FSim",
    "Rn", "GroupedPauli", "Power", "BarrierGate", "FSIM", "ParamCustomGate", "TwoParamCustomGate",
    "gene_univ_parameterized_gate", "gene_univ_two_params_gate",
]
