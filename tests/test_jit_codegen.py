"""The per-circuit code generator: its kernels compile for every gate kind (CPU only).

``mqa_circuit_jit_compile`` builds the forward or back-sweep schedule, emits one HIP kernel per
pass and compiles it with hipRTC, as the library does in the background for large launches.
No GPU is needed; parity of the compiled kernels is checked in ``test_gpu_jit.py``.
"""
import ctypes

import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.core import gates as G
from tests.conftest import hea, random_circuit


def compile_circuit(circ, n, dtype_code, adjoint):
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    nc = ir.NativeCircuit(circ, names)
    sec = ctypes.c_double(0)
    _native.check(_native.lib().mqa_circuit_jit_compile(nc.handle, n, dtype_code, int(adjoint), ctypes.byref(sec)))
    return sec.value


@pytest.mark.parametrize("dtype_code", [_native.MQA_C64, _native.MQA_C128])
@pytest.mark.parametrize("adjoint", [False, True])
def test_every_gate_kind_compiles(dtype_code, adjoint):
    """Random circuit over every kind, local and non-local controls, 3-qubit matrices."""
    n = 9
    circ = random_circuit(n, 60, 5, ctrl_prob=0.3)
    rng = np.random.default_rng(1)
    m4 = np.linalg.qr(rng.normal(size=(16, 16)) + 1j * rng.normal(size=(16, 16)))[0]
    circ += G.UnivMathGate("u4", m4).on([0, 3, 5, 8])
    circ += G.RY("q").on(2, [7, 8])
    compile_circuit(circ, n, dtype_code, adjoint)


@pytest.mark.parametrize("dtype_code", [_native.MQA_C64, _native.MQA_C128])
@pytest.mark.parametrize("mode", [2, 3])
def test_phase_free_forms_compile(dtype_code, mode):
    """The phase-free forward (2) and back-sweep (3) kernels of the hermitian gradient: shear
    rotations, one-amplitude RZ phases, dropped GP, CNOTs across the tile as selects."""
    n = 14
    circ = hea(n, 2, encoder=True)
    circ += G.GlobalPhase("g").on(3)
    circ += G.RZ("c").on(2, 13)
    compile_circuit(circ, n, dtype_code, mode)


def test_hea_compiles_both_directions():
    circ = hea(14, 3, encoder=True)
    assert compile_circuit(circ, 14, _native.MQA_C64, False) >= 0
    assert compile_circuit(circ, 14, _native.MQA_C64, True) >= 0


def test_bad_arguments_are_reported():
    circ = mq.Circuit().h(3)
    nc = ir.NativeCircuit(circ, [])
    st = _native.lib().mqa_circuit_jit_compile(nc.handle, 2, _native.MQA_C64, 0, None)
    assert st == _native.MQA_ERR_ARG


def test_removed_persistent_form_is_rejected():
    circ = hea(8, 1)
    nc = ir.NativeCircuit(circ, list(circ.params_name))
    for mode in (4, 5, 7):
        assert _native.lib().mqa_circuit_jit_compile(nc.handle, 8, _native.MQA_C64, mode, None) == _native.MQA_ERR_ARG


@pytest.mark.parametrize("mode", [9, 11])
def test_single_buffer_back_sweep_compiles(mode):
    """Single-buffer complex64 back-sweep kernels (exact 9 and phase-free 11)."""
    circ = hea(14, 2, encoder=True)
    circ += G.RY("c").on(2, 13)
    circ += G.X.on(12, [0, 13])
    compile_circuit(circ, 14, _native.MQA_C64, mode)


def test_single_buffer_rejects_dense_blocks_and_forward_plans():
    rng = np.random.default_rng(4)
    m = np.linalg.qr(rng.normal(size=(8, 8)) + 1j * rng.normal(size=(8, 8)))[0]
    circ = hea(8, 1) + G.UnivMathGate("u3", m).on([0, 1, 2])
    nc = ir.NativeCircuit(circ, list(circ.params_name))
    assert _native.lib().mqa_circuit_jit_compile(nc.handle, 8, _native.MQA_C64, 9, None) == _native.MQA_ERR_ARG
    plain = hea(8, 1)
    nc2 = ir.NativeCircuit(plain, list(plain.params_name))
    assert _native.lib().mqa_circuit_jit_compile(nc2.handle, 8, _native.MQA_C64, 8, None) == _native.MQA_ERR_ARG


def test_single_buffer_back_sweep_compiles_c128():
    circ = hea(12, 2, encoder=True)
    circ += G.RY("c").on(2, 11)
    compile_circuit(circ, 12, _native.MQA_C128, 11)


@pytest.mark.parametrize("mode", [0, 1, 2, 3])
def test_coefficients_come_ready_made(tmp_path, monkeypatch, mode):
    """Compiled passes read each parameterised op's coefficients from the coefficient table
    (shear t and s, the doubled angle's cos and sin, or the (cos, sin) pair) instead of deriving
    them in every lane; back-sweep inverses negate them in the read."""
    monkeypatch.setenv("MQAMD_JIT_DUMP", str(tmp_path))
    circ = hea(12, 2, encoder=True)
    circ += G.RZ("c").on(2, 11)  # controlled: exact RZ in both plan forms
    compile_circuit(circ, 12, _native.MQA_C64, mode)
    src = "".join(p.read_text() for p in sorted(tmp_path.glob("pass*.hip")))
    assert src.count("coef_row[") >= 2 * len(circ.params_name)
    assert "ang_row[" not in src and "shear_of(" not in src and "double_angle(" not in src
    assert ("-coef_row[" in src) == bool(mode & 1)
    assert ("Shear<T> h{" in src) == bool(mode & 2)
