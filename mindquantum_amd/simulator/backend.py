"""The MI355X state-vector backend: Python host side of the C ABI.

Mirrors mindquantum/simulator/mqsim.py (MQSim): argument validation, the ``version``
logic of get_expectation_with_grad (encoder / ansatz / both), and the output slicing
``f = res[:, :, 0]``, ``g = res[:, :, 1:]``.  All numerical work runs in libmqamd.so; there
is no CPU code path.
"""
import ctypes
import numbers

import numpy as np

from .. import _native
from .. import dtype as mqdtype
from ..core.circuit import Circuit
from ..core import gates as G
from ..core.gates import BasicGate
from ..core.operators import Hamiltonian
from ..ir import NativeCache, resolve
from .utils import GradOpsWrapper, check_parallel_worker


def _check_encoder(data, n):
    """type_value_check.py:24-34"""
    if not isinstance(data, np.ndarray):
        raise ValueError(f"encoder parameters need numpy array, but get {type(data)}")
    if data.ndim != 2:
        raise ValueError("encoder data requires a two dimension numpy array")
    if data.shape[1] != n:
        raise ValueError(f"encoder parameters size do not match with encoder parameters name, need {n} but get "
                         f"{data.shape[1]}.")


def _check_ansatz(data, n):
    """type_value_check.py:37-48"""
    if not isinstance(data, np.ndarray):
        raise ValueError(f"ansatz parameters need numpy array, but get {type(data)}")
    if data.ndim != 1:
        raise ValueError("ansatz data requires a one dimension numpy array")
    if data.shape[0] != n:
        raise ValueError(f"ansatz parameters size do not match with ansatz parameters name, need {n} but get "
                         f"{data.shape[0]}")


class MQVectorAMD:
    """State-vector simulator on one MI355X (gfx950) GPU."""

    name = "mqvector_amd"

    def __init__(self, name, n_qubits, seed=None, dtype=None):
        if not isinstance(n_qubits, (int, np.integer)) or n_qubits < 1:
            raise ValueError(f"n_qubits should be a positive int, but get {n_qubits}")
        self.name = name
        self.n_qubits = int(n_qubits)
        self.seed = int(np.random.randint(1, 2**23)) if seed is None else int(seed)
        self.dtype = mqdtype.to_mq_type(dtype)
        self._code = mqdtype.native_code(self.dtype)
        self._circ_cache = NativeCache(64)
        h = ctypes.c_void_p()
        _native.check(_native.lib().mqa_sim_create(self.n_qubits, self._code, self.seed & 0xFFFFFFFF, ctypes.byref(h)))
        self._h = h

    @classmethod
    def _wrap(cls, name, handle, n_qubits, seed, dtype):
        obj = cls.__new__(cls)
        obj.name = name
        obj.n_qubits = n_qubits
        obj.seed = seed
        obj.dtype = mqdtype.to_mq_type(dtype)
        obj._code = mqdtype.native_code(obj.dtype)
        obj._circ_cache = NativeCache(64)
        obj._h = handle
        return obj

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and _native._lib is not None:
            _native._lib.mqa_sim_destroy(h)
            self._h = None

    def __getstate__(self):  # mqsim.py:84-98: the state itself is not pickled
        return {"name": self.name, "n_qubits": self.n_qubits, "seed": self.seed, "dtype": self.dtype}

    def __setstate__(self, st):
        self.__init__(st["name"], st["n_qubits"], st["seed"], st["dtype"])

    # ---- circuits
    def _native_circuit(self, circ, names):
        return self._circ_cache.get(circ, names)

    def _check_circ(self, circ):
        if not isinstance(circ, Circuit):
            raise TypeError(f"circuit requires a Circuit, but get {type(circ)}")
        if circ.n_qubits > self.n_qubits:
            raise ValueError(f"Circuit has {circ.n_qubits} qubits, which is more than simulator qubits "
                             f"{self.n_qubits}.")

    def _check_ham(self, ham, i=0):
        if not isinstance(ham, Hamiltonian):
            raise TypeError(f"hams's element requires a Hamiltonian, but get {type(ham)}")
        if not mqdtype.is_same_precision(self.dtype, ham.dtype):
            raise TypeError(
                f"Data type of {self.name} simulator is {mqdtype.precision_str(self.dtype)} ({self.dtype.__name__}), "
                f"but {i}th hamiltonian is {mqdtype.precision_str(ham.dtype)} ({ham.dtype.__name__}).")
        if ham.n_qubits > self.n_qubits or (ham.how_to == 2 and ham.n_qubits != self.n_qubits):
            raise ValueError(f"Hamiltonian qubits is {ham.n_qubits}, not match with simulator qubits number "
                             f"{self.n_qubits}")

    def apply_gate(self, gate, pr=None, diff=False):
        """Apply ``gate`` (mqsim.py:129-160).  With ``diff`` a parameterised gate whose parameters
        require gradients applies the derivative of its matrix with respect to the angle value on
        the control subspace and zeroes every amplitude outside it (vector_state.tpp:165-437 with
        diff; rot_pauli.cpp:431-446 SetToZeroExcept): lowered here to the derivative as a dense
        matrix with the gate's controls, then diag(0, 1) on every control qubit.  Gates without
        gradient-requiring parameters apply their matrix, as in the reference."""
        if not isinstance(gate, BasicGate):
            raise TypeError(f"gate requires a quantum gate, but get {type(gate)}")
        if not gate.obj_qubits:
            raise ValueError(f"Gate {gate.name} has not been applied on any qubit; use .on()")
        if diff and getattr(gate, "parameterized", False):
            if isinstance(gate, G.MultiParamsGate) or not hasattr(gate, "diff_matrix"):
                # vector_state.tpp:328-366 names U3, Rn, FSim and two-parameter custom gates as not
                # applicable in differential form
                raise ValueError(f"Can not apply differential format of {gate.name} gate on quantum states "
                                 "currently.")
            if gate.coeff.requires_grad_parameters:
                if pr is None:
                    raise ValueError("apply a parameterized gate needs a parameter_resolver")
                dm = gate.diff_matrix(pr)
                circ = Circuit([G.UnivMathGate(f"d{gate.name}", dm).on(gate.obj_qubits, gate.ctrl_qubits)])
                keep_one = np.diag([0.0, 1.0]).astype(np.complex128)
                for c in gate.ctrl_qubits:
                    circ += G.UnivMathGate("P1", keep_one).on(c)
                return self.apply_circuit(circ)
        return self.apply_circuit(Circuit([gate]), pr)

    def apply_circuit(self, circuit, pr=None):
        self._check_circ(circuit)
        if len(circuit) == 0:
            return None
        if circuit.has_measure_gate:
            raise ValueError("apply_circuit here runs unitary circuits; sample measurements with sampling()")
        names, vals = resolve(circuit, pr)
        nc = self._native_circuit(circuit, names)
        nc.set_custom(vals[None, :])
        _native.check(_native.lib().mqa_sim_apply_circuit(self._h, nc.handle, _native.ptr(vals), len(vals)))
        return None

    def sampling(self, circuit, pr=None, shots=1, seed=None):
        """Sample a measurement-ending circuit (mqsim.py:456-491, vector_state.tpp:1326-1373): the
        non-measurement gates run on a copy of the state, the device draws amplitude indices by
        cumulative probability for sorted uniforms from numpy's generator seeded with ``seed``, and
        the measured qubits' bits are collected per key.  Agrees with the reference in distribution
        (its mt19937 stream is not reproduced)."""
        from ..core.measurement import MeasureResult

        self._check_circ(circuit)
        measures = circuit.all_measures
        if not measures:
            raise ValueError("circuit must have at least one measurement gate.")
        if circuit.n_qubits > self.n_qubits:
            raise ValueError(f"Circuit has {circuit.n_qubits} qubits, which is more than simulator qubits.")
        if not isinstance(shots, (int, np.integer)) or isinstance(shots, bool):
            raise TypeError(f"sampling shots requires a int, but get {type(shots)}")
        if shots < 1:
            raise ValueError(f"sampling shots should be not less than 1, but get {shots}")
        if not circuit.is_measure_end:
            raise ValueError("only measurement-ending circuits are sampled here (every measurement after the last "
                             "gate on its qubit)")
        keys = [m.key for m in measures]
        if len(set(keys)) != len(keys):
            raise ValueError("measure keys must be unique")
        body = circuit.remove_measure()
        if seed is None:
            seed = int(np.random.randint(1, 2 << 20))
        u = np.sort(np.random.default_rng(seed).random(int(shots)))
        idx = np.zeros(int(shots), dtype=np.uint64)
        handle, vals = None, np.zeros(0)
        if len(body):
            names, vals = resolve(body, pr)
            nc = self._native_circuit(body, names)
            nc.set_custom(vals[None, :])
            handle = nc.handle
        _native.check(_native.lib().mqa_sim_sample(self._h, handle, _native.ptr(vals) if len(vals) else None,
                                                    len(vals), _native.ptr(u), int(shots), _native.ptr(idx)))
        res = MeasureResult()
        res.add_measure(measures)
        km = res.keys_map
        samples = np.zeros((int(shots), len(keys)), dtype=np.int64)
        for m in measures:
            samples[:, km[m.key]] = ((idx >> np.uint64(m.obj_qubits[0])) & np.uint64(1)).astype(np.int64)
        res.collect_data(samples)
        return res

    def apply_hamiltonian(self, hamiltonian):
        self._check_ham(hamiltonian)
        _native.check(_native.lib().mqa_sim_apply_hamiltonian(self._h, hamiltonian.native()))

    # ---- state access
    def get_qs(self, ket=False):
        out = np.empty(1 << self.n_qubits, dtype=self.dtype)
        _native.check(_native.lib().mqa_sim_get_qs(self._h, _native.ptr(out), out.size))
        if ket:
            from ..utils import ket_string

            return "\n".join(ket_string(out))
        return out

    def set_qs(self, quantum_state):
        qs = np.asarray(quantum_state)
        if qs.ndim != 1 or qs.size != (1 << self.n_qubits):
            raise ValueError(f"Quantum state should be a 1-D array of length {1 << self.n_qubits}")
        norm = np.sqrt(np.sum(np.abs(qs) ** 2))
        if norm < 1e-12:
            raise ValueError("Norm of quantum state can not be zero.")
        qs = np.ascontiguousarray(qs / norm, dtype=self.dtype)  # mqsim.py:537-544
        _native.check(_native.lib().mqa_sim_set_qs(self._h, _native.ptr(qs), qs.size))

    def reset(self):
        _native.check(_native.lib().mqa_sim_reset(self._h))

    def copy(self):
        h = ctypes.c_void_p()
        _native.check(_native.lib().mqa_sim_copy(self._h, ctypes.byref(h)))
        return MQVectorAMD._wrap(self.name, h, self.n_qubits, self.seed, self.dtype)

    def astype(self, dtype, seed=None):
        code = mqdtype.native_code(dtype)
        h = ctypes.c_void_p()
        _native.check(_native.lib().mqa_sim_astype(self._h, code, ctypes.byref(h)))
        return MQVectorAMD._wrap(self.name, h, self.n_qubits, self.seed if seed is None else seed, dtype)

    # ---- expectations
    def get_expectation(self, hamiltonian, circ_right=None, circ_left=None, simulator_left=None, pr=None):
        self._check_ham(hamiltonian)
        if circ_right is None:
            circ_right = Circuit()
        self._check_circ(circ_right)
        if circ_left is not None:
            self._check_circ(circ_left)
        if simulator_left is not None:
            if not isinstance(simulator_left, MQVectorAMD):
                raise TypeError("simulator_left should be a simulator of the same backend")
            if simulator_left.n_qubits != self.n_qubits or simulator_left.dtype != self.dtype:
                raise ValueError("simulator_left should have the same n_qubits and dtype as this simulator")
        names = list(circ_right.params_name)
        if circ_left is not None:
            names += [n for n in circ_left.params_name if n not in names]
        vals = np.zeros(len(names), dtype=np.float64)
        if names:
            if pr is None:
                raise ValueError("Parameterized circuit needs a parameter resolver.")
            mapping = dict(pr.items()) if hasattr(pr, "items") else dict(zip(names, np.asarray(pr).ravel()))
            vals = np.array([float(np.real(mapping[n])) for n in names], dtype=np.float64)
        ncs = [self._native_circuit(c, names) for c in (circ_right, circ_left) if c is not None and len(c)]
        for nc in ncs:
            nc.set_custom(vals[None, :])
        cr = self._native_circuit(circ_right, names).handle if len(circ_right) else None
        cl = self._native_circuit(circ_left, names).handle if circ_left is not None and len(circ_left) else None
        left = simulator_left._h if simulator_left is not None else None
        out = np.zeros(2, dtype=np.float64)
        if circ_left is not None and cl is None:
            # an empty left circuit still makes the bra the (left) initial state
            cl_empty = self._native_circuit(Circuit(), names)
            cl = cl_empty.handle
        _native.check(_native.lib().mqa_sim_expectation_lr(self._h, left, hamiltonian.native(), cr, cl,
                                                           _native.ptr(vals), len(vals), _native.ptr(out)))
        return complex(out[0], out[1])

    def get_expectation_with_grad(self, hams, circ_right, circ_left=None, simulator_left=None,
                                  parallel_worker=None, pr_shift=False):
        if isinstance(hams, Hamiltonian):
            hams = [hams]
        elif not isinstance(hams, list):
            raise TypeError(f"hams requires a Hamiltonian or a list of Hamiltonian, but get {type(hams)}")
        for i, h in enumerate(hams):
            self._check_ham(h, i)
        self._check_circ(circ_right)
        non_hermitian = False
        if circ_left is not None:
            self._check_circ(circ_left)
            non_hermitian = True
        if simulator_left is not None:
            if not isinstance(simulator_left, type(self)):
                raise TypeError("simulator_left should be a simulator of the same backend")
            if simulator_left.n_qubits != self.n_qubits:
                raise ValueError("simulator_left should have the same n_qubits as this simulator, "
                                 f"which is {self.n_qubits}, but get {simulator_left.n_qubits}")
            if simulator_left.dtype != self.dtype:
                raise ValueError("simulator_left should have the same data type as this simulator "
                                 f"(aka {self.dtype}), but get {simulator_left.dtype}.")
            non_hermitian = True
        if non_hermitian and simulator_left is None:
            simulator_left = self
        if circ_left is None:
            circ_left = circ_right
        if pr_shift and non_hermitian:  # mqsim.py:372-388: the non-hermitian gradient takes precedence
            pr_shift = False
        if circ_right.has_measure_gate:
            raise ValueError("circuit for variational algorithm cannot have measure gate")
        check_parallel_worker(parallel_worker)
        ansatz_params_name = list(circ_right.all_ansatz.keys())
        encoder_params_name = list(circ_right.all_encoder.keys())
        if non_hermitian:  # mqsim.py:323-331: the left circuit's names follow the right circuit's
            ansatz_params_name += [n for n in circ_left.all_ansatz.keys() if n not in ansatz_params_name]
            encoder_params_name += [n for n in circ_left.all_encoder.keys() if n not in encoder_params_name]
        if set(ansatz_params_name) & set(encoder_params_name):
            raise RuntimeError("Parameter cannot be both encoder and ansatz parameter.")
        version = "both"
        if not ansatz_params_name:
            version = "encoder"
        if not encoder_params_name:
            version = "ansatz"
        names = encoder_params_name + ansatz_params_name
        if max(circ_left.n_qubits, circ_right.n_qubits) > self.n_qubits:
            raise ValueError(f"Simulator has {self.n_qubits} qubits, but circuit has "
                             f"{max(circ_left.n_qubits, circ_right.n_qubits)} qubits.")
        nc = self._native_circuit(circ_right, names)
        ncl = self._native_circuit(circ_left, names) if non_hermitian else None
        ham_handles = (ctypes.c_void_p * len(hams))(*[h.native().value for h in hams])
        n_enc, n_ans = len(encoder_params_name), len(ansatz_params_name)

        def grad_ops(*inputs_):
            inputs = [np.array(x) if isinstance(x, list) else x for x in inputs_]
            if version == "both" and len(inputs) != 2:
                raise ValueError("Need two inputs!")
            if version in ("encoder", "ansatz") and len(inputs) != 1:
                raise ValueError("Need one input!")
            if version in ("both", "encoder"):
                if inputs[0].ndim == 1:
                    inputs[0] = np.array([inputs[0]])
                _check_encoder(inputs[0], n_enc)
                enc = np.ascontiguousarray(inputs[0], dtype=np.float64)
                ans = (np.ascontiguousarray(inputs[1], dtype=np.float64) if version == "both"
                       else np.zeros(0, dtype=np.float64))
                if version == "both":
                    _check_ansatz(inputs[1], n_ans)
            else:
                _check_ansatz(inputs[0], n_ans)
                enc = np.zeros((1, 0), dtype=np.float64)
                ans = np.ascontiguousarray(inputs[0], dtype=np.float64)
            batch = enc.shape[0]
            if nc.custom or (ncl is not None and ncl.custom):  # per-sample matrices of custom gates
                rows = np.concatenate([enc, np.broadcast_to(ans, (batch, ans.size))], axis=1)
                nc.set_custom(rows)
                if ncl is not None:
                    ncl.set_custom(rows)
            if pr_shift:
                from .param_shift import gradient as shift_gradient

                fc, gc = shift_gradient(self, hams, circ_right, encoder_params_name, ansatz_params_name, enc, ans)
                fc, gc = fc.astype(self.dtype), gc.astype(self.dtype)
                if version == "both":
                    return fc, gc[:, :, :n_enc], gc[:, :, n_enc:]
                return fc, gc
            f = np.zeros((batch, len(hams), 2), dtype=np.float64)
            if non_hermitian:
                g = np.zeros((batch, len(hams), n_enc + n_ans, 2), dtype=np.float64)
                _native.check(_native.lib().mqa_sim_grad_nonherm(
                    self._h, simulator_left._h, ham_handles, len(hams), nc.handle, ncl.handle,
                    _native.ptr(enc) if n_enc else None, batch, n_enc, _native.ptr(ans) if n_ans else None, n_ans,
                    _native.ptr(f), _native.ptr(g)))
                gc = (g[..., 0] + 1j * g[..., 1]).astype(self.dtype)
            else:
                g = np.zeros((batch, len(hams), n_enc + n_ans), dtype=np.float64)
                _native.check(_native.lib().mqa_sim_grad_batch(
                    self._h, ham_handles, len(hams), nc.handle, _native.ptr(enc) if n_enc else None, batch, n_enc,
                    _native.ptr(ans) if n_ans else None, n_ans, _native.ptr(f), _native.ptr(g)))
                gc = g.astype(self.dtype)
            fc = (f[..., 0] + 1j * f[..., 1]).astype(self.dtype)
            if version == "both":
                return fc, gc[:, :, :n_enc], gc[:, :, n_enc:]
            return fc, gc

        wrapper = GradOpsWrapper(grad_ops, hams, circ_right, circ_left, encoder_params_name, ansatz_params_name,
                                 parallel_worker, self)
        s = f"{self.n_qubits} qubit" + ("" if self.n_qubits == 1 else "s") + f" {self.name} VQA Operator"
        wrapper.set_str(s)
        wrapper.native_circuits = [nc] + ([ncl] if ncl is not None else [])  # handles the autotuner adjusts
        return wrapper


def inner_product(bra, ket):
    """<bra|ket> of two simulators of the same backend (mq_blas.py / blas.h InnerProduct)."""
    a = bra.backend if hasattr(bra, "backend") else bra
    b = ket.backend if hasattr(ket, "backend") else ket
    if a.n_qubits != b.n_qubits:
        raise ValueError("two states must have the same qubit number")
    if a.dtype != b.dtype:
        raise TypeError("two states must have the same dtype")
    out = np.zeros(2, dtype=np.float64)
    _native.check(_native.lib().mqa_inner_product(a._h, b._h, _native.ptr(out)))
    return complex(out[0], out[1])


def is_number(x):
    return isinstance(x, numbers.Number)
