"""Circuit: an ordered list of gates with parameter bookkeeping.

Host-side mirror of mindquantum/core/circuit/circuit.py for the state-vector path: gate
insertion (+=, +, append), the parameter name views (params_name, encoder_params_name,
ansatz_params_name, all_encoder, all_ansatz), as_encoder / as_ansatz / no_grad, hermitian()
(reversed order, each gate hermitian-ed, circuit.py:831-845) and the gate helper methods.
"""
import copy

from . import gates as G
from .mutation import bump, epoch
from .parameterresolver import ParameterResolver


class _NameList(list):
    """Ordered parameter names; ``.keys()`` mirrors the reference's ordered-dict views."""

    def keys(self):
        return list(self)


class Circuit(list):
    """Ordered list of gates."""

    def __init__(self, gates=None):
        super().__init__()
        if gates is None:
            return
        if isinstance(gates, G.BasicGate):
            gates = [gates]
        for g in gates:
            self.append(g)

    # ---- construction (every list mutation advances the mutation epoch)
    def append(self, gate):
        if isinstance(gate, Circuit):
            for g in gate:
                super().append(g)
            bump()
            return self
        if not isinstance(gate, G.BasicGate):
            raise TypeError(f"Circuit only accepts gates, but get {type(gate)}")
        if not gate.obj_qubits:
            raise ValueError(f"Gate {gate.name} has not been applied on any qubit; use .on()")
        super().append(gate)
        bump()
        return self

    def extend(self, gates):
        for g in gates:
            self.append(g)
        return self

    def insert(self, index, gate):
        if not isinstance(gate, G.BasicGate):
            raise TypeError("Circuit only accepts gates")
        super().insert(index, gate)
        bump()

    def __setitem__(self, index, value):
        items = list(value) if isinstance(index, slice) else [value]
        if not all(isinstance(g, G.BasicGate) for g in items):
            raise TypeError("Circuit only accepts gates")
        super().__setitem__(index, value)
        bump()

    def __delitem__(self, index):
        super().__delitem__(index)
        bump()

    def pop(self, index=-1):
        g = super().pop(index)
        bump()
        return g

    def remove(self, gate):
        super().remove(gate)
        bump()

    def clear(self):
        super().clear()
        bump()

    def reverse(self):
        super().reverse()
        bump()

    def sort(self, *args, **kwargs):
        super().sort(*args, **kwargs)
        bump()

    def __imul__(self, n):
        res = super().__imul__(n)
        bump()
        return res

    def __iadd__(self, other):
        if isinstance(other, (Circuit, list)):
            for g in other:
                self.append(g)
        else:
            self.append(other)
        return self

    def __add__(self, other):
        out = Circuit(list(self))
        out += other
        return out

    def __radd__(self, other):
        out = Circuit(other if isinstance(other, (list, Circuit)) else [other])
        out += self
        return out

    def __getitem__(self, item):
        res = super().__getitem__(item)
        if isinstance(item, slice):
            return Circuit(res)
        return res

    def __copy__(self):
        return Circuit(copy.deepcopy(list(self)))

    # ---- properties
    @property
    def n_qubits(self):
        qs = [q for g in self for q in g.obj_qubits + g.ctrl_qubits]
        return max(qs) + 1 if qs else 0

    def _names(self, which, pred):
        # the views are kept until the mutation epoch advances (any change to any circuit, gate or
        # parameter resolver): batched calls read them for every circuit on every call
        now = epoch()
        views = self.__dict__.get("_name_views")
        if views is None or views[0] != now:
            views = (now, {})
            self._name_views = views
        names = views[1].get(which)
        if names is None:
            names, seen = [], set()
            for g in self:
                for pr in g.param_resolvers():
                    for k in pr.keys():
                        if k not in seen and pred(pr, k):
                            seen.add(k)
                            names.append(k)
            views[1][which] = names
        return _NameList(names)

    @property
    def params_name(self):
        return self._names(0, lambda pr, k: True)

    @property
    def encoder_params_name(self):
        return self._names(1, lambda pr, k: k in pr.encoder_parameters)

    @property
    def ansatz_params_name(self):
        return self._names(2, lambda pr, k: k not in pr.encoder_parameters)

    @property
    def all_encoder(self):
        return self.encoder_params_name

    @property
    def all_ansatz(self):
        return self.ansatz_params_name

    @property
    def parameterized(self):
        return any(getattr(g, "parameterized", False) for g in self)

    @property
    def all_measures(self):
        """Measurement gates in circuit order."""
        from .measurement import Measure

        return [g for g in self if isinstance(g, Measure)]

    @property
    def has_measure_gate(self):
        return bool(self.all_measures)

    @property
    def is_measure_end(self):
        """Every measurement comes after the last other gate acting on its qubit (and each qubit is
        measured at most once)."""
        from .measurement import Measure

        done = set()
        for g in self:
            qs = set(g.obj_qubits) | set(g.ctrl_qubits)
            if qs & done:
                return False
            if isinstance(g, Measure):
                done |= qs
        return True

    def remove_measure(self):
        from .measurement import Measure

        return Circuit([g for g in self if not isinstance(g, Measure)])

    def measure(self, key, obj=None):
        """Append ``Measure(key).on(obj)`` (``measure(q)``: key ``q<q>``)."""
        from .measurement import Measure

        if obj is None:
            self.append(Measure().on(key))
        else:
            self.append(Measure(key).on(obj))
        return self

    @property
    def is_noise_circuit(self):
        return False

    # ---- parameter flags
    def _each_pr(self, fn):
        for g in self:
            for pr in g.param_resolvers():
                fn(pr)

    def as_encoder(self, inplace=True):
        circ = self if inplace else copy.deepcopy(self)
        circ._each_pr(lambda pr: pr.as_encoder())
        return circ

    def as_ansatz(self, inplace=True):
        circ = self if inplace else copy.deepcopy(self)
        circ._each_pr(lambda pr: pr.as_ansatz())
        return circ

    def no_grad(self):
        self._each_pr(lambda pr: pr.no_grad())
        return self

    def requires_grad(self):
        self._each_pr(lambda pr: pr.requires_grad())
        return self

    def hermitian(self):
        """Reversed circuit with every gate hermitian-ed (circuit.py:831-845)."""
        return Circuit([g.hermitian() for g in reversed(self)])

    def get_cpp_obj(self, hermitian=False):  # pragma: no cover - reference naming only
        raise NotImplementedError("the HIP backend compiles circuits with mindquantum_amd.ir")

    # ---- helpers (circuit.py gate shortcuts)
    def _add(self, gate, obj, ctrl):
        self.append(gate.on(obj, ctrl))
        return self

    def x(self, obj, ctrl=None):
        return self._add(G.X, obj, ctrl)

    def y(self, obj, ctrl=None):
        return self._add(G.Y, obj, ctrl)

    def z(self, obj, ctrl=None):
        return self._add(G.Z, obj, ctrl)

    def h(self, obj, ctrl=None):
        return self._add(G.H, obj, ctrl)

    def s(self, obj, ctrl=None):
        return self._add(G.S, obj, ctrl)

    def t(self, obj, ctrl=None):
        return self._add(G.T, obj, ctrl)

    def swap(self, obj, ctrl=None):
        return self._add(G.SWAP, obj, ctrl)

    def rx(self, para, obj, ctrl=None):
        return self._add(G.RX(para), obj, ctrl)

    def ry(self, para, obj, ctrl=None):
        return self._add(G.RY(para), obj, ctrl)

    def rz(self, para, obj, ctrl=None):
        return self._add(G.RZ(para), obj, ctrl)

    def rxx(self, para, obj, ctrl=None):
        return self._add(G.Rxx(para), obj, ctrl)

    def ryy(self, para, obj, ctrl=None):
        return self._add(G.Ryy(para), obj, ctrl)

    def rzz(self, para, obj, ctrl=None):
        return self._add(G.Rzz(para), obj, ctrl)

    def phase_shift(self, para, obj, ctrl=None):
        return self._add(G.PhaseShift(para), obj, ctrl)

    def __repr__(self):
        return "Circuit(" + ", ".join(repr(g) for g in self) + ")"


def UN(gate, maps_obj, maps_ctrl=None):  # pylint: disable=invalid-name
    """Apply ``gate`` on every qubit of ``maps_obj`` (core/circuit/module_circuit.py UN)."""
    if isinstance(maps_obj, int):
        maps_obj = list(range(maps_obj))
    circ = Circuit()
    for i, q in enumerate(maps_obj):
        ctrl = None if maps_ctrl is None else maps_ctrl[i]
        circ += gate.on(q, ctrl)
    return circ
