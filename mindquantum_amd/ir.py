"""Lowering of a Circuit to the gate IR of include/mqamd.h.

A circuit is compiled against an ordered parameter-name list (encoder names first, then
ansatz names, gradient.h:149-158): every parameterised gate gets ``angle = const +
sum coeff * params[index]`` terms, fixed gates their kind or a dense matrix.  Composite gates
(U3, FSim, Givens, SWAPalpha, Rxy/Rxz/Ryz, RotPauliString, GroupedPauli) are expanded into
their exact lowering first (``BasicGate.lowered``), so gate indices refer to that expansion.
"""
import collections
import ctypes
import hashlib
import weakref

import numpy as np

from . import _native
from .core import gates as G
from .core.mutation import epoch
from .core.parameterresolver import ParameterResolver


def _is_custom(g):
    return getattr(g, "ir_kind", None) == "PMATRIX"


def lower(circ, names, layout=None):
    """Return (gates, term_param, term_coeff, mats, n_mat_entries, requires_grad_per_param) numpy arrays.

    Parameterised matrix gates (custom gates, parameterised Rn) become one MQA_GATE_PMATRIX record
    plus one MQA_GATE_PMATRIX_TAP record per further parameter expression; ``layout`` (a list, if
    given) receives (gate, offset, n_targets, n_expressions) for each of them, the offset (in doubles)
    of its block [G][G^dagger][M_0][M_1]... in a sample's row of the custom-matrix table."""
    index = {n: i for i, n in enumerate(names)}
    prims = [p for g in circ for p in g.lowered()]
    records = []  # (gate, ir kind, parameter resolver or None, custom offset)
    offset = 0
    for g in prims:
        if _is_custom(g):
            k = len(g.obj_qubits)
            if k > 2:
                raise ValueError(f"parameterised matrix gate {g.name} acts on {k} qubits; the HIP backend runs them "
                                 "on one or two target qubits")
            prs = g.param_resolvers()
            dd = 2 << (2 * k)
            if layout is not None:
                layout.append((g, offset, k, len(prs)))
            records.append((g, "PMATRIX", prs[0], offset))
            for j in range(1, len(prs)):
                records.append((g, "PMATRIX_TAP", prs[j], offset + (2 + j) * dd))
            offset += (2 + len(prs)) * dd
        else:
            records.append((g, None, None, -1))
    gates = np.zeros(len(records), dtype=_native.GATE_DTYPE)
    t_param, t_coeff, mats = [], [], []
    no_grad = set()

    def terms_of(rec, pr):
        rec["const_val"] = float(np.real(pr.const))
        rec["term_begin"] = len(t_param)
        for name, c in pr.items():
            if name not in index:
                raise ValueError(f"Parameter {name} is not in the parameter list")
            t_param.append(index[name])
            t_coeff.append(float(np.real(c)))
        rec["n_term"] = len(pr)
        no_grad.update(pr.no_grad_parameters)
        rec["requires_grad"] = 1 if pr.requires_grad_parameters else 0

    for i, (g, custom_kind, pr, off) in enumerate(records):
        rec = gates[i]
        rec["n_obj"] = len(g.obj_qubits)
        for k, q in enumerate(g.obj_qubits):
            rec["obj"][k] = q
        rec["ctrl_mask"] = sum(1 << c for c in g.ctrl_qubits)
        if custom_kind is not None:
            rec["kind"] = _native.GATE_KIND[custom_kind]
            rec["mat_begin"] = off
            terms_of(rec, pr)
        elif isinstance(g, G.ParameterGate):
            rec["kind"] = _native.GATE_KIND[g.ir_kind]
            terms_of(rec, g.coeff)
        elif isinstance(g, G.NoneParameterGate):
            kind = g.ir_kind
            rec["kind"] = _native.GATE_KIND[kind]
            if kind == "MATRIX":
                m = g.matrix()
                rec["mat_begin"] = len(mats)
                mats.extend(np.asarray(m, dtype=np.complex128).ravel())
        else:
            raise ValueError(f"Gate {g.name} is not supported by the HIP backend")
    mats_arr = np.zeros(2 * max(len(mats), 1), dtype=np.float64)
    if mats:
        m = np.asarray(mats, dtype=np.complex128)
        mats_arr[0::2] = m.real
        mats_arr[1::2] = m.imag
    rg = np.array([0 if n in no_grad else 1 for n in names], dtype=np.uint8)
    return (gates, np.asarray(t_param, dtype=np.int32), np.asarray(t_coeff, dtype=np.float64), mats_arr,
            len(mats), rg)


def custom_table(layout, names, values):
    """Rows of the custom-matrix table (mqa_circuit_set_custom_matrices) for parameter values
    ``values`` [rows, len(names)]: per gate the matrix G, G^dagger and M_j = dG/dv_j G^dagger of each
    parameter expression v_j, evaluated on the host (the reference evaluates its numba functions per
    sample, vector_state.tpp:402-413)."""
    values = np.atleast_2d(np.asarray(values, dtype=np.float64))
    rows = values.shape[0]
    stride = max(off + (2 + m) * (2 << (2 * k)) for _, off, k, m in layout)
    table = np.zeros((rows, stride), dtype=np.float64)
    index = {n: i for i, n in enumerate(names)}
    for r in range(rows):
        pr = {n: values[r, i] for n, i in index.items()}
        for g, off, k, m in layout:
            gm = np.asarray(g.matrix(pr), dtype=np.complex128)
            blocks = [gm, gm.conj().T] + [np.asarray(d, dtype=np.complex128) @ gm.conj().T
                                          for d in g.diff_matrices(pr)]
            flat = np.concatenate([b.ravel() for b in blocks])
            table[r, off:off + 2 * flat.size:2] = flat.real
            table[r, off + 1:off + 2 * flat.size:2] = flat.imag
    return table, stride


def digest(arrays):
    """Content digest of lowered arrays (gate records, angle terms, matrices, no-grad flags)."""
    gates, tp, tc, mats, n_mats, rg = arrays
    h = hashlib.blake2b(digest_size=16)
    for a in (gates, tp, tc, mats, rg):
        h.update(np.ascontiguousarray(a).tobytes())
    h.update(int(n_mats).to_bytes(8, "little"))
    return h.digest()


class NativeCircuit:
    """C ABI circuit handle for (circuit, parameter order)."""

    def __init__(self, circ, names, arrays=None, layout=None):
        self.names = list(names)
        if arrays is None:
            layout = []
            arrays = lower(circ, self.names, layout)
        self.arrays = arrays
        self.custom = layout or []  # (gate, offset, targets, expressions) of parameterised matrix gates
        self.digest = digest(self.arrays)
        gates, tp, tc, mats, n_mats, rg = self.arrays
        h = ctypes.c_void_p()
        rgp = _native.ptr(rg) if len(rg) else None
        _native.check(_native.lib().mqa_circuit_create(
            _native.ptr(gates), len(gates), _native.ptr(tp) if len(tp) else None,
            _native.ptr(tc) if len(tc) else None, len(tp), _native.ptr(mats), n_mats, len(self.names), rgp,
            ctypes.byref(h)))
        self.handle = h
        # tile widths chosen by the autotuner for this circuit: {(dtype code, adjoint): qubits}
        for (code, adjoint), k in getattr(circ, "_mqa_tiles", {}).items():
            self.set_tile_qubits(code, adjoint, k)

    def set_custom(self, values):
        """Evaluate the parameterised matrix gates for the parameter rows ``values`` [rows, len(names)]
        (one row: shared by every sample) and hand them to the library for the calls that follow."""
        if not self.custom:
            return
        table, stride = custom_table(self.custom, self.names, values)
        table = np.ascontiguousarray(table)
        _native.check(_native.lib().mqa_circuit_set_custom_matrices(self.handle, table.shape[0], stride,
                                                                     _native.ptr(table)))

    def set_tile_qubits(self, dtype_code, adjoint, tile_qubits):
        """Per-circuit tile width of the forward (adjoint False) or back-sweep passes; negative
        restores the global option."""
        _native.check(_native.lib().mqa_circuit_set_tile_qubits(self.handle, int(dtype_code), int(bool(adjoint)),
                                                                 int(tile_qubits)))

    def tile_qubits(self, dtype_code, adjoint):
        k = ctypes.c_int32(0)
        _native.check(_native.lib().mqa_circuit_get_tile_qubits(self.handle, int(dtype_code), int(bool(adjoint)),
                                                                 ctypes.byref(k)))
        return k.value

    def schedule(self, n_qubits, dtype_code, adjoint=False):
        """Tile-pass schedule: list of (local qubit mask, [gate indices])."""
        n = ctypes.c_int64(0)
        _native.check(_native.lib().mqa_circuit_schedule(self.handle, n_qubits, dtype_code, int(adjoint), None, 0,
                                                          ctypes.byref(n)))
        buf = np.zeros(n.value, dtype=np.int64)
        _native.check(_native.lib().mqa_circuit_schedule(self.handle, n_qubits, dtype_code, int(adjoint),
                                                          _native.ptr(buf), n.value, ctypes.byref(n)))
        out, pos = [], 1
        for _ in range(int(buf[0])):
            mask, cnt = int(buf[pos]), int(buf[pos + 1])
            out.append((mask, [int(x) for x in buf[pos + 2: pos + 2 + cnt]]))
            pos += 2 + cnt
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        if h is not None and _native._lib is not None:
            _native._lib.mqa_circuit_destroy(h)
            self.handle = None


class NativeCache:
    """Native circuits of Circuit objects, reused across calls.

    The reference lowers its circuit on every call (mqsim.py: ``get_cpp_obj()``).  Here a native
    circuit (and with it its schedules, device plans and compiled passes) is kept per (circuit
    object, parameter order): while the global mutation epoch (mindquantum_amd.core.mutation) is
    unchanged the circuit cannot have changed and the entry is returned at once; after any change
    anywhere the circuit is lowered again and the entry reused only if the lowered content is
    identical (a gate replaced, a parameter made no-grad, ... build a new one).  Entries are held
    weakly by circuit (dropped when the circuit object dies) and at most ``maxsize`` are kept
    (least recently used first out).
    """

    def __init__(self, maxsize):
        self.maxsize = int(maxsize)
        self._d = collections.OrderedDict()

    def __len__(self):
        return len(self._d)

    def get(self, circ, names):
        key = (id(circ), tuple(names))
        hit = self._d.get(key)
        now = epoch()
        if hit is not None and hit[0]() is circ:
            if hit[2] == now:
                self._d.move_to_end(key)
                return hit[1]
            layout = []
            arrays = lower(circ, names, layout)
            if digest(arrays) == hit[1].digest:
                hit[1].custom = layout  # the same content: take the current gate objects
                self._d[key] = (hit[0], hit[1], now)
                self._d.move_to_end(key)
                return hit[1]
            nc = NativeCircuit(circ, names, arrays, layout)
        else:
            nc = NativeCircuit(circ, names)
        ref = weakref.ref(circ, lambda _r, k=key, d=self._d: d.pop(k, None) if k in d and d[k][0] is _r else None)
        self._d[key] = (ref, nc, now)
        self._d.move_to_end(key)
        while len(self._d) > self.maxsize:
            self._d.popitem(last=False)
        return nc


def param_names(circ):
    """(encoder names, ansatz names) in order of appearance."""
    return list(circ.encoder_params_name), list(circ.ansatz_params_name)


def resolve(circ, pr):
    """Parameter vector of a circuit for ``pr`` (dict, ParameterResolver, array or None)."""
    names = list(circ.params_name)
    if not names:
        return names, np.zeros(0, dtype=np.float64)
    if pr is None:
        raise ValueError("Parameterized circuit needs a parameter resolver.")
    if isinstance(pr, ParameterResolver):
        vals = dict(pr.items())
    elif isinstance(pr, dict):
        vals = pr
    else:
        arr = np.asarray(pr, dtype=np.float64).ravel()
        if arr.size != len(names):
            raise ValueError(f"Need {len(names)} parameters, but get {arr.size}.")
        return names, arr
    missing = [n for n in names if n not in vals]
    if missing:
        raise ValueError(f"Parameters {missing} are not given.")
    return names, np.array([float(np.real(vals[n])) for n in names], dtype=np.float64)
