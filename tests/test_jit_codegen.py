"""The per-circuit code generator: its kernels compile for every gate kind (CPU only).

``mqa_circuit_jit_compile`` builds the forward or back-sweep schedule, emits one HIP kernel per
pass and compiles it with hipRTC, as the library does in the background for large launches.
No GPU is needed; parity of the compiled kernels is checked in ``test_gpu_jit.py``.
"""
import contextlib
import ctypes
import re

import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from mindquantum_amd.core import gates as G
from tests.conftest import hea, random_circuit


@contextlib.contextmanager
def wave_local_relays(value):
    saved = _native.get_option("jit_wave_local")
    _native.set_option("jit_wave_local", value)
    try:
        yield
    finally:
        _native.set_option("jit_wave_local", saved)


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


@pytest.mark.parametrize("mode", [5, 7])
def test_persistent_back_sweep_compiles(mode):
    """Persistent complex64 back-sweep kernels (exact 5 and phase-free 7)."""
    circ = hea(14, 2, encoder=True)
    circ += G.RY("c").on(2, 13)
    compile_circuit(circ, 14, _native.MQA_C64, mode)


def test_persistent_is_c64_back_sweep_only():
    circ = hea(8, 1)
    nc = ir.NativeCircuit(circ, list(circ.params_name))
    for dtype_code, mode in ((_native.MQA_C128, 5), (_native.MQA_C64, 4)):
        assert _native.lib().mqa_circuit_jit_compile(nc.handle, 8, dtype_code, mode, None) == _native.MQA_ERR_ARG


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


def test_single_buffer_transitions_without_workgroup_barriers(tmp_path, monkeypatch):
    """A back-sweep re-lay between two register layouts that agree on the positions the wave index
    takes (the top K - 10 non-register positions) orders one wave's own LDS accesses only; the
    others keep workgroup barriers."""
    monkeypatch.setenv("MQAMD_JIT_DUMP", str(tmp_path))
    with wave_local_relays(1):
        compile_circuit(hea(20, 2, encoder=True), 20, _native.MQA_C64, 11)  # 12-qubit tiles: 4 waves
    local = barrier = 0
    for path in tmp_path.glob("pass*.hip"):
        src = path.read_text()
        pos = [set(int(v) for v in m.split(","))
               for m in re.findall(r"constexpr uint8_t pos_\d+\[4\] = \{([^}]*)\}", src)]
        parts = src.split("  {  // group ")  # parts[g + 1] opens with group g; its tail is the re-lay
        for g in range(1, len(pos)):
            relay = parts[g].rsplit("    }\n", 1)[-1]
            if sorted(set(range(12)) - pos[g - 1])[-2:] == sorted(set(range(12)) - pos[g])[-2:]:
                assert relay.count("wave_sync();") == 4 and "__syncthreads" not in relay
                local += 1
            else:
                assert relay.count("__syncthreads();") == 4 and "wave_sync" not in relay
                barrier += 1
    assert local > 0 and barrier > 0


def test_forward_group_ends_without_workgroup_barrier_when_wave_local(tmp_path, monkeypatch):
    """A forward pass's register group ends with a wave-scope fence instead of a workgroup barrier
    when the next group keeps every wave's tile indices; the last group keeps the barrier before
    the tile goes back to HBM."""
    monkeypatch.setenv("MQAMD_JIT_DUMP", str(tmp_path))
    with wave_local_relays(1):
        compile_circuit(hea(20, 2, encoder=True), 20, _native.MQA_C64, 2)  # phase-free forward, 12-qubit tiles
    local = barrier = 0
    for path in tmp_path.glob("pass*.hip"):
        src = path.read_text()
        pos = [set(int(v) for v in m.split(","))
               for m in re.findall(r"constexpr uint8_t pos\[4\] = \{([^}]*)\}", src)]
        ends = re.findall(r"group_store<T, false>\(f, a0, off, p, l\);\n    (\w+)\(\);", src)
        assert len(ends) == len(pos) and ends[-1] == "__syncthreads"
        for g in range(len(pos) - 1):
            same = sorted(set(range(12)) - pos[g])[-2:] == sorted(set(range(12)) - pos[g + 1])[-2:]
            assert ends[g] == ("wave_sync" if same else "__syncthreads")
            local += same
            barrier += not same
    assert local > 0 and barrier > 0


def test_wave_local_relays_are_off_by_default(tmp_path, monkeypatch):
    """Option jit_wave_local = 0 (the default): every re-lay takes workgroup barriers."""
    monkeypatch.setenv("MQAMD_JIT_DUMP", str(tmp_path))
    assert _native.get_option("jit_wave_local") == 0
    for mode in (2, 11):
        compile_circuit(hea(20, 2, encoder=True), 20, _native.MQA_C64, mode)
        assert not any("wave_sync();" in p.read_text() for p in tmp_path.glob("pass*.hip"))
