"""Autotuner of the tile width (qubits per LDS tile) of a circuit's passes.

Wider tiles hold more gates per pass (fewer HBM sweeps) but take more LDS per block (fewer
resident blocks to overlap memory and arithmetic); the best width depends on the circuit, the
precision and whether the pass carries the back-sweep's second state.  The tuner times the
candidate widths on the device with the circuit itself and keeps the fastest: it is stored on the
Circuit (``circ._mqa_tiles``), so every simulator that lowers the circuit afterwards uses it, and
it is set on the given gradient operator's handles directly.  Compiled passes are built before
timing (each candidate's first run is untimed), so only the steady state is compared.  The widths
the circuit has when the search starts are kept unless a candidate is faster by more than
``tolerance`` (a fraction of the time): candidates a timing's noise apart would otherwise be chosen
at random from run to run.
"""
import time

import numpy as np

from . import _native
from . import dtype as mqdtype

C64_CANDIDATES = (11, 12, 13)
C128_CANDIDATES = (10, 11, 12)


def _defaults(code):
    return C64_CANDIDATES if code == _native.MQA_C64 else C128_CANDIDATES


def _store(circ, code, adjoint, k):
    tiles = dict(getattr(circ, "_mqa_tiles", {}))
    tiles[(code, bool(adjoint))] = int(k)
    circ._mqa_tiles = tiles


def _pick(times, start, tolerance):
    """The fastest key of ``times``, unless ``start`` is within ``tolerance`` of it."""
    best = min(times, key=times.get)
    if start in times and times[start] <= times[best] * (1.0 + tolerance):
        return start
    return best


def _timed(fn, repeats):
    fn()  # plans, compiled passes, allocations
    best = float("inf")
    for _ in range(repeats):
        t0 = time.perf_counter()
        fn()
        best = min(best, time.perf_counter() - t0)
    return best


def tune_apply(circ, n_qubits, dtype=None, pr=None, candidates=None, repeats=1, tolerance=0.02):
    """Time ``apply_circuit(circ, pr)`` on a scratch simulator for every candidate width of the
    forward passes; keep the fastest.  Returns {width: seconds}."""
    from .simulator import Simulator

    dt = mqdtype.to_mq_type(dtype)
    code = mqdtype.native_code(dt)
    sim = Simulator("mqvector_amd", n_qubits, dtype=dt)
    names = list(circ.params_name)
    nc = sim.backend._native_circuit(circ, names)
    start = nc.tile_qubits(code, False)
    times = {}
    jit = _native.get_option("jit")
    _native.set_option("jit", 2)
    try:
        for k in candidates or _defaults(code):
            nc.set_tile_qubits(code, False, k)

            def run():
                sim.reset()
                sim.apply_circuit(circ, pr)

            times[k] = _timed(run, repeats)
    finally:
        _native.set_option("jit", jit)
    best = _pick(times, start, tolerance)
    _store(circ, code, False, best)
    nc.set_tile_qubits(code, False, best)
    return times


def tune_grad(grad_ops, *inputs, fwd_candidates=None, adj_candidates=None, repeats=1,
              tolerance=0.02):
    """Coordinate search over the back-sweep width, then the forward width, of a gradient
    operator (``Simulator.get_expectation_with_grad(...)``) on its own inputs.  The choice is set
    on the operator and stored on its circuit.  Returns {(forward, back-sweep): seconds}."""
    handles = getattr(grad_ops, "native_circuits", None)
    if not handles:
        raise ValueError("grad_ops does not come from a mindquantum_amd simulator")
    code = mqdtype.native_code(grad_ops.sim.dtype)
    fwd_c = list(fwd_candidates or _defaults(code))
    adj_c = list(adj_candidates or _defaults(code))
    circs = [grad_ops.circ_right] + ([grad_ops.circ_left] if len(handles) > 1 else [])

    def apply(fk, ak):
        for nc in handles:
            nc.set_tile_qubits(code, False, fk)
            nc.set_tile_qubits(code, True, ak)

    times = {}
    jit = _native.get_option("jit")
    _native.set_option("jit", 2)
    try:
        fk = handles[0].tile_qubits(code, False)
        start = (fk, handles[0].tile_qubits(code, True))
        for ak in adj_c:
            apply(fk, ak)
            times[(fk, ak)] = _timed(lambda: grad_ops(*inputs), repeats)
        ak = _pick(times, start, tolerance)[1]
        for f in fwd_c:
            if (f, ak) not in times:
                apply(f, ak)
                times[(f, ak)] = _timed(lambda: grad_ops(*inputs), repeats)
    finally:
        _native.set_option("jit", jit)
    best_f, best_a = _pick(times, start, tolerance)
    apply(best_f, best_a)
    for c in circs:
        _store(c, code, False, best_f)
        _store(c, code, True, best_a)
    return times
