"""CPU checks of the C ABI library: symbols, loud failure without a GPU, and the tile-pass
scheduler (host-only entry point mqa_circuit_schedule) - every gate exactly once, locality
of every pass, dependency order, and equivalence of the reordered gate sequence."""
import os
import re

import numpy as np
import pytest

import mindquantum_amd as mq
from mindquantum_amd import _native, ir
from oracle import dense
from tests.conftest import ROOT, hea, params_for, random_circuit


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "mqamd.h")).read()
    return sorted(set(re.findall(r"^(?:int|const char\*)\s+(mqa_\w+)\(", txt, flags=re.M)))


def test_library_exports_every_declared_symbol():
    lib = _native.lib()
    declared = _header_functions()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), name
    assert declared == _native.exported_symbols()


def test_version_and_options():
    assert b"gfx950" in _native.lib().mqa_version()
    old = _native.get_option("tile_qubits_c64")
    _native.set_option("tile_qubits_c64", 10)
    assert _native.get_option("tile_qubits_c64") == 10
    _native.set_option("tile_qubits_c64", -1)
    assert _native.get_option("tile_qubits_c64") == old
    with pytest.raises(ValueError):
        _native.set_option("no_such_option", 1)


@pytest.mark.skipif(_native.device_count() > 0, reason="checks the no-GPU error path")
def test_sim_create_fails_loudly_without_gpu():
    with pytest.raises(_native.NativeError, match="no HIP device"):
        mq.Simulator("mqvector_amd", 4)


def test_bad_arguments_raise():
    from mindquantum_amd.core import gates as G

    with pytest.raises(ValueError):
        G.X.on([0, 0])
    with pytest.raises(ValueError):
        G.RX("a").on(0, 0)
    circ = mq.Circuit().rx("a", 0)
    with pytest.raises(ValueError):
        ir.NativeCircuit(circ, [])  # parameter missing from the list


def _check_schedule(circ, sched, n_eff, K, adjoint):
    gates = [i for i, g in enumerate(circ) if not (hasattr(g, "ir_kind") and g.ir_kind == "I")]
    order = [gi for _, ops in sched for gi in ops]
    assert sorted(order) == sorted(gates)
    pos = {gi: k for k, gi in enumerate(order)}
    for mask, ops in sched:
        assert bin(mask).count("1") == min(K, n_eff)
        for gi in ops:
            for q in circ[gi].obj_qubits:
                assert (mask >> q) & 1, "target outside the pass's local qubits"
    # gates sharing a qubit keep their relative order (reversed in the back-sweep)
    touched = [set(g.obj_qubits + g.ctrl_qubits) for g in circ]
    for a in gates:
        for b in gates:
            if a < b and touched[a] & touched[b]:
                assert (pos[a] < pos[b]) != adjoint


@pytest.mark.parametrize("n,depth,seed", [(8, 200, 1), (12, 300, 2), (20, 400, 3), (28, 300, 4)])
def test_schedule_is_a_valid_reordering(n, depth, seed):
    circ = random_circuit(n, depth, seed)
    names = list(circ.params_name)
    nc = ir.NativeCircuit(circ, names)
    for adjoint in (False, True):
        for dt in (0, 1):
            K = _native.get_option("adj_tile_qubits_c64" if adjoint else "tile_qubits_c64") if dt == 0 else \
                _native.get_option("adj_tile_qubits_c128" if adjoint else "tile_qubits_c128")
            sched = nc.schedule(n, dt, adjoint)
            _check_schedule(circ, sched, max(n, 6), K, adjoint)


def test_schedule_fuses_hea_layers():
    circ = hea(20, 8)
    nc = ir.NativeCircuit(circ, list(circ.params_name))
    sched = nc.schedule(20, 0)
    assert len(sched) * 10 < len(circ)  # >10x fewer state sweeps than gates


def test_unfused_schedule_option():
    circ = hea(10, 2)
    _native.set_option("max_ops_per_pass", 1)
    try:
        nc = ir.NativeCircuit(circ, list(circ.params_name))
        sched = nc.schedule(10, 0)
    finally:
        _native.set_option("max_ops_per_pass", -1)
    assert len(sched) == len(circ)


def test_cost_model_partition_splits_the_heavy_pass():
    """Without an explicit cap the scheduler keeps the cheapest of a few capped partitions: on the
    20-qubit encoder + HEA-8 gradient circuit it splits the greedy schedule's first pass (all
    layers on the low qubits) without adding passes."""
    from mindquantum_amd.algorithm import hardware_efficient_ansatz
    from mindquantum_amd.algorithm.builders import encoder_rx

    circ = encoder_rx(20) + hardware_efficient_ansatz(20, 8)
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    for adjoint in (False, True):
        auto = ir.NativeCircuit(circ, names).schedule(20, 0, adjoint)
        _native.set_option("max_ops_per_pass", 100000)  # an explicit cap that never binds: greedy
        try:
            greedy = ir.NativeCircuit(circ, names).schedule(20, 0, adjoint)
        finally:
            _native.set_option("max_ops_per_pass", -1)
        _check_schedule(circ, auto, 20, 12, adjoint)
        assert len(auto) <= len(greedy)
        assert max(len(ops) for _, ops in auto) < max(len(ops) for _, ops in greedy)


@pytest.mark.parametrize("seed", range(4))
def test_scheduled_order_gives_the_same_state(seed):
    """Apply the gates in the scheduler's order with the dense simulation: same state."""
    n = 8
    _native.set_option("tile_qubits_c64", 6)
    _native.set_option("lowq_c64", 1)
    try:
        circ = random_circuit(n, 80, 50 + seed)
        names, vals = params_for(circ, seed)
        values = dict(zip(names, vals))
        nc = ir.NativeCircuit(circ, names)
        sched = nc.schedule(n, 0)
    finally:
        _native.set_option("tile_qubits_c64", -1)
        _native.set_option("lowq_c64", -1)
    assert len(sched) > 3
    order = [gi for _, ops in sched for gi in ops]
    reordered = mq.Circuit([circ[i] for i in order])
    a = dense.state(circ, values, n)
    b = dense.state(reordered, values, n)
    assert np.allclose(a, b, atol=1e-12)


def test_per_circuit_tile_width():
    """Per-circuit tile widths (set by the autotuner) reach the scheduler and travel with the
    Circuit object to every handle lowered from it; out-of-range widths are rejected."""
    from mindquantum_amd.algorithm import hardware_efficient_ansatz

    circ = hardware_efficient_ansatz(16, 2)
    nc = ir.NativeCircuit(circ, list(circ.params_name))
    default = nc.tile_qubits(_native.MQA_C64, False)
    assert default == _native.get_option("tile_qubits_c64")
    wide = len(nc.schedule(16, _native.MQA_C64))
    nc.set_tile_qubits(_native.MQA_C64, False, 8)
    assert nc.tile_qubits(_native.MQA_C64, False) == 8
    assert len(nc.schedule(16, _native.MQA_C64)) > wide  # narrower tiles: more passes
    assert nc.tile_qubits(_native.MQA_C128, False) == _native.get_option("tile_qubits_c128")
    nc.set_tile_qubits(_native.MQA_C64, False, -1)
    assert nc.tile_qubits(_native.MQA_C64, False) == default
    with pytest.raises(ValueError):
        nc.set_tile_qubits(_native.MQA_C64, True, 14)
    circ._mqa_tiles = {(_native.MQA_C64, True): 10}
    nc2 = ir.NativeCircuit(circ, list(circ.params_name))
    assert nc2.tile_qubits(_native.MQA_C64, True) == 10
