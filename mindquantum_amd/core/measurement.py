"""Measurement gate and sampling result (mindquantum/core/gates/measurement.py).

``Measure(key).on(q)`` marks qubit q for measurement; ``Simulator.sampling`` runs a
measurement-ending circuit (every measurement after the last gate on its qubit) and returns a
``MeasureResult`` whose keys are stored in reverse circuit order, so that a bit string reads the
last measurement first (measurement.py:254-298).
"""
import copy

import numpy as np

from .gates import BasicGate, _qubits


class Measure(BasicGate):
    """Computational-basis measurement of one qubit; the key defaults to ``q<qubit>``."""

    def __init__(self, name="", reset_to=None):
        super().__init__("M", 1)
        if reset_to not in (None, 0, 1):
            raise ValueError("reset_to should be None, 0 or 1")
        self.key = name
        self.reset_to = reset_to

    def on(self, obj_qubits, ctrl_qubits=None):
        if ctrl_qubits:
            raise ValueError("Measure gate can not have control qubit")
        obj = _qubits(obj_qubits)
        if len(obj) != 1:
            raise ValueError("Measure gate only apply on a single qubit")
        new = copy.deepcopy(self)
        new.obj_qubits = obj
        new.ctrl_qubits = []
        if not new.key:
            new.key = f"q{obj[0]}"
        return new

    def lowered(self):
        raise ValueError("a measurement gate can not run in apply_circuit / expectation here; use "
                         "Simulator.sampling for measurement-ending circuits")

    def hermitian(self):
        raise ValueError("measurement gate is not unitary")

    def __repr__(self):
        return f"M({self.key})" + (f"({self.obj_qubits[0]})" if self.obj_qubits else "")


class MeasureResult:
    """Sampled bit strings of a set of measurement keys."""

    def __init__(self):
        self.measures = []
        self._keys = []
        self._samples = np.zeros((0, 0), dtype=np.int64)
        self.bit_string_data = {}
        self.shots = 0

    @property
    def keys(self):
        return list(self._keys)

    def add_measure(self, measure):
        if isinstance(measure, Measure):
            measure = [measure]
        for meas in measure:
            if not isinstance(meas, Measure):
                raise ValueError("Measurement gates need to be objects of class 'Measurement' ")
        for meas in reversed(list(measure)):
            if meas.key in self._keys:
                raise ValueError(f"Measure key {meas.key} already defined.")
            self.measures.append(meas)
            self._keys.append(meas.key)

    @property
    def keys_map(self):
        return {k: i for i, k in enumerate(self._keys)}

    def collect_data(self, samples):
        samples = np.asarray(samples)
        self._samples = samples
        self.shots = len(samples)
        out = {}
        for row in samples:
            s = "".join(str(int(b)) for b in row)
            out[s] = out.get(s, 0) + 1
        self.bit_string_data = {k: out[k] for k in sorted(out)}

    @property
    def samples(self):
        return self._samples

    @property
    def data(self):
        return self.bit_string_data

    def __repr__(self):
        return f"MeasureResult(shots={self.shots}, keys={self._keys}, data={self.bit_string_data})"
