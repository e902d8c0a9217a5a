#!/usr/bin/env python3
"""Benchmark of the MI355X state-vector backend.

Headline (``value``): batched adjoint gradients per second - 20 qubits, batch 1024, complex64:
encoder RX(x_i) on every qubit (x ~ N(0,1)), 8-layer hardware-efficient ansatz (RY, RZ on every
qubit, CNOT ladder i -> i+1; 320 parameters ~ U[0, 2pi)), H = sum_i w_i Z_i Z_{i+1 mod 20} +
sum_i v_i X_i with w, v ~ U(-1, 1) (40 terms).  One step = one forward + backward of the
torch.autograd operator (MQOps) over the whole batch; the adjoint sweep runs in the forward.

Also reported in the same JSON line, each with its error against the complex128 oracle (or, where no
oracle runs at that size, the property the mathematics fixes): the same workload in complex128; the
16-qubit HEA-10 complex128 single-vector gradient with the CPU oracle beside it; 28-qubit gate
application (layered random circuit, depth 200, complex64 / complex128) in HBM GB/s, fused (tile
passes) and unfused (one gate per pass), with the CPU oracle on a 2-layer sample; the uneven batch
of 512 MaxCut QAOA circuits (12..24 qubits, p = 4, complex64, one mqa_grad_items call); one 32-qubit
complex64 and one 31-qubit complex128 gradient (HEA-4, 64 random Pauli strings, 4 checkpoints) and
their 24-qubit analog beside the CPU oracle.  ``roofline`` is the dominant kernel of the headline (the
adjoint tile pass) timed with HIP events on the library's stream; ``cpu_baseline`` times the CPU
oracle (a restatement of the reference CPU mqvector path) on 64 samples of the headline workload in
the reference's three thread layouts on this host's cores.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# HBM bytes per launch of the dominant kernel, measured with FETCH_SIZE / WRITE_SIZE counter runs
# scaled by a copy kernel of known size (the summary file names the commands and the factors)
TRAFFIC_FILE = os.path.join(ROOT, "profiles", "adjoint_20q_b1024_c64_hbm_counters.md")
METRIC = "adjoint grads/s (20q, batch 1024, c64 vs c128); 28q gate-apply HBM GB/s"
# sampled rows of the headline batch checked against the complex128 oracle
ERR_ROWS = np.array([0, 1, 2, 3, 97, 255, 256, 400, 511, 512, 640, 777, 900, 1000, 1022, 1023])


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None


def log(msg):
    print(msg, file=sys.stderr, flush=True)


MARKS = False


def mark(name):
    """With --mark-workloads, log the start of a workload on the host's monotonic and boot-time clocks
    (nanoseconds), so that a kernel trace taken of the run can be split by workload."""
    if MARKS:
        log(f"workload {name} begins mono_ns={time.monotonic_ns()} "
            f"boot_ns={time.clock_gettime_ns(time.CLOCK_BOOTTIME)}")


def host_cores():
    cores = len(os.sched_getaffinity(0))
    return min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))


def coeff_sum(op):
    return float(sum(abs(complex(v)) for v in op.terms.values()))


def err_stats(f, g, f_ref, g_ref, csum):
    """Error against a complex128 reference: |f - f_ref| / max(1, sum |c|) and
    max |g - g_ref| / max |g_ref|."""
    f_err = float(np.max(np.abs(np.asarray(f) - np.asarray(f_ref))))
    g_err = float(np.max(np.abs(np.asarray(g) - np.asarray(g_ref))))
    g_max = float(np.max(np.abs(g_ref)))
    return {"f_abs": f_err, "f_rel_sum_c": f_err / max(1.0, csum), "g_abs": g_err, "g_rel_max_g": g_err / g_max}


def oracle_grad(circ, n, op, enc, ans, threads):
    from mindquantum_amd import ir
    import mindquantum_amd as mq
    from oracle import cpu

    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    return cpu.grad_batch(n, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()], enc, ans,
                          batch_threads=threads, worker_threads=1)


# ---------------------------------------------------------------- headline
def run_adjoint(dtype, n, layers, batch, steps, warmup, seed, profile, tune=False):
    import torch

    import mindquantum_amd as mq
    from mindquantum_amd import _native, autotune
    from mindquantum_amd.algorithm import batched_adjoint
    from mindquantum_amd.framework import MQOps

    circ, op, enc, ans = batched_adjoint(n, layers, batch, seed)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)
    tiles = None
    if tune:  # untimed: tile widths of the forward and back-sweep passes, chosen on this workload
        c64 = dtype == mq.complex64
        times = autotune.tune_grad(grad_ops, enc, ans, fwd_candidates=(12, 13) if c64 else (11, 12),
                                   adj_candidates=(11, 12) if c64 else (10, 11, 12))
        code = mq.dtype.native_code(dtype)
        nc = grad_ops.native_circuits[0]
        tiles = {"forward_back_sweep": [nc.tile_qubits(code, False), nc.tile_qubits(code, True)],
                 "seconds": {f"{f}/{a}": round(t, 4) for (f, a), t in times.items()}}
        log(f"autotune ({np.dtype(dtype).name}): {tiles}")
    ops = MQOps(grad_ops)
    x = torch.tensor(enc, dtype=torch.float32, requires_grad=True)
    a = torch.nn.Parameter(torch.tensor(ans, dtype=torch.float32))

    def step():
        out = ops(x, a)
        out.sum().backward()
        return out

    for i in range(warmup):
        step()
        log(f"warmup step {i + 1}/{warmup} done ({np.dtype(dtype).name})")
    if profile:
        _native.profile_reset()
        _native.set_option("profile", 1)
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        log(f"timed step {i + 1}/{steps} done ({np.dtype(dtype).name})")
    dt = time.perf_counter() - t0
    stats = {}
    if profile:
        _native.set_option("profile", 0)
        for k in ("tile_pass_fwd", "tile_pass_adj", "hamiltonian"):
            stats[k] = _native.profile(k)
    # the values the timed steps computed on the sampled rows (the float32 inputs MQOps passes on)
    rows = ERR_ROWS[ERR_ROWS < batch]
    x32 = x.detach().numpy().astype(np.float64)
    a32 = a.detach().numpy().astype(np.float64)
    f, g_enc, g_ans = grad_ops(x32[rows], a32)
    sample = {"rows": rows, "enc": x32[rows], "ans": a32, "f": f[:, 0], "g": np.concatenate([g_enc[:, 0],
                                                                                          g_ans[:, 0]], 1).real,
              "circ": circ, "op": op}
    return {"grads_per_s": batch * steps / dt, "ms_per_step": dt / steps * 1e3, "stats": stats, "tiles": tiles,
            "n_params": len(circ.params_name), "n_gates": len(circ), "sample": sample}


def headline_errors(samples, n):
    """Errors of the sampled rows of each precision against the complex128 oracle on the same
    (float32-rounded) inputs."""
    s0 = samples[0]
    t0 = time.perf_counter()
    f_ref, g_ref = oracle_grad(s0["circ"], n, s0["op"], s0["enc"], s0["ans"], min(len(s0["rows"]), host_cores()))
    log(f"c128 oracle on {len(s0['rows'])} sampled rows: {time.perf_counter() - t0:.1f} s")
    csum = coeff_sum(s0["op"])
    out = []
    for s in samples:
        e = err_stats(s["f"], s["g"], f_ref[:, 0], g_ref[:, 0].real, csum)
        e["rows"] = int(len(s["rows"]))
        out.append(e)
    return out


# ---------------------------------------------------------------- single-vector 16q HEA-10 c128
def run_hea16(reps, cpu):
    """The CPU-runnable single-vector gradient (b = 1, complex128) on the GPU, beside the CPU oracle
    (which the reference runs inline with OpenMP over all cores, gradient.h:271-273)."""
    import mindquantum_amd as mq
    from mindquantum_amd.algorithm import hea_open_chain

    from mindquantum_amd import _native

    circ, op, vals = hea_open_chain(16, 10, 0)
    sim = mq.Simulator("mqvector_amd", 16, dtype=mq.complex128)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, mq.complex128), circ)

    def timed_calls():
        for _ in range(3):
            grad_ops(vals)
        t0 = time.perf_counter()
        for _ in range(reps):
            out = grad_ops(vals)
        return (time.perf_counter() - t0) / reps * 1e3, out

    # interpreted passes: what the first calls of a plan run while its kernels compile
    jit_prev = _native.get_option("jit")
    _native.set_option("jit", 0)
    try:
        interp_ms, _ = timed_calls()
    finally:
        _native.set_option("jit", -1)
    # steady state: an untimed call waits for the per-circuit compiled passes (a recurring plan
    # compiles in the background by default), then the timed calls run them with the default options
    c0 = _native.get_option("jit_compile_ms")
    _native.set_option("jit", 2)
    _native.set_option("jit_min_amps", 0)
    try:
        grad_ops(vals)
    finally:
        _native.set_option("jit", -1)
        _native.set_option("jit_min_amps", -1)
    compile_s = (_native.get_option("jit_compile_ms") - c0) / 1e3
    launches0 = _native.get_option("jit_launches")
    gpu_ms, (f, g) = timed_calls()
    assert _native.get_option("jit_launches") > launches0
    _native.set_option("jit", jit_prev)  # the later workloads run with the caller's setting again
    res = {"qubits": 16, "layers": 10, "params": len(vals), "terms": len(op.terms), "dtype": "complex128",
           "gpu_ms_per_grad": round(gpu_ms, 3), "calls_timed": reps, "compiled_passes": True,
           "compile_seconds_untimed": round(compile_s, 1), "gpu_ms_per_grad_interpreted": round(interp_ms, 3)}
    if cpu:
        from oracle import cpu as ocpu
        from mindquantum_amd import ir

        cores = host_cores()
        ocpu.set_num_threads(cores)
        names = list(circ.params_name)
        t0 = time.perf_counter()
        f_ref, g_ref = ocpu.grad_batch(16, np.complex128, ir.lower(circ, names), [mq.Hamiltonian(op).terms_ir()],
                                       np.zeros((1, 0)), vals, batch_threads=1)
        cpu_ms = (time.perf_counter() - t0) * 1e3
        res.update({"cpu_ms_per_grad": round(cpu_ms, 1), "cpu_cores": cores, "cpu_kind": "port",
                    "speedup_vs_cpu": round(cpu_ms / gpu_ms, 1),
                    "err_vs_c128_oracle": err_stats(f[0, 0], g[0, 0].real, f_ref[0, 0], g_ref[0, 0].real,
                                                    coeff_sum(op))})
    log(f"16q HEA-10 c128 single vector: {res}")
    return res


# ---------------------------------------------------------------- 28q gate application
def run_gate_apply(dtype, n, depth, seed, unfused_depth, tune=False, variants=True):
    import mindquantum_amd as mq
    from mindquantum_amd import _native, autotune
    from mindquantum_amd.algorithm import layered_random_circuit

    amp = 8 if dtype == mq.complex64 else 16
    circ = layered_random_circuit(n, depth, seed)
    tuned = None
    if tune:  # untimed: the tile width of the passes, chosen on this circuit
        times = autotune.tune_apply(circ, n, dtype)
        tuned = {"tile_qubits": circ._mqa_tiles[(mq.dtype.native_code(dtype), False)],
                 "seconds": {k: round(v, 4) for k, v in times.items()}}
        log(f"autotune apply {n}q {np.dtype(dtype).name}: {tuned}")
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
    # untimed run of the same circuit: plan, allocations and the compiled passes (jit = 2)
    c0 = _native.get_option("jit_compile_ms")
    sim.apply_circuit(circ)
    compile_s = (_native.get_option("jit_compile_ms") - c0) / 1e3
    # three timed applications from |0>; the best is reported, with the kernel figures of that application
    reps, times, prof = 3, [], []
    _native.set_option("profile", 1)
    for _ in range(reps):
        sim.reset()
        _native.profile_reset()
        t0 = time.perf_counter()
        sim.apply_circuit(circ)
        times.append(time.perf_counter() - t0)
        prof.append(_native.profile("tile_pass_fwd"))
    _native.set_option("profile", 0)
    best = min(range(reps), key=lambda i: times[i])
    dt = times[best]
    launches, ms, byts = prof[best]
    passes = launches
    state = (1 << n) * amp
    norm = mq.inner_product(sim, sim).real
    res = {
        "gates": len(circ), "passes": passes, "seconds": round(dt, 4),
        "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
        "frac_of_peak": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "gate_equivalent_gbs": round(len(circ) * 2 * state / dt / 1e9, 1),
        "kernel_ms": round(ms, 3), "launches": launches,
        "compiled_passes": _native.get_option("jit") != 0, "compile_seconds_untimed": round(compile_s, 1),
        "autotune": tuned, "norm_drift": abs(norm - 1.0), "seconds_each": [round(t, 4) for t in times],
    }
    log(f"gate apply {n}q {np.dtype(dtype).name}: {res}")
    if not variants:
        return res, sim
    # unfused: one gate per pass, on a shorter prefix of the same circuit.  Single-gate passes
    # stream memory; they run on the interpreted kernel, with the tile width that streams fastest
    # (timed here over the widths 10..12; the reference's schedule of one sweep per gate)
    short = circ[: unfused_depth * (n + n // 2)]
    _native.set_option("max_ops_per_pass", 1)
    jit = _native.get_option("jit")
    _native.set_option("jit", 0)
    key = "tile_qubits_c64" if dtype == mq.complex64 else "tile_qubits_c128"
    per_k = {}
    try:
        sim2 = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
        for k in ((10, 11, 12) if tune else (12,)):
            _native.set_option(key, k)
            sim2.apply_circuit(short[:4])
            _native.profile_reset()
            _native.set_option("profile", 1)
            t0 = time.perf_counter()
            sim2.apply_circuit(short)
            dt2 = time.perf_counter() - t0
            _native.set_option("profile", 0)
            launches2, ms2, byts2 = _native.profile("tile_pass_fwd")
            per_k[k] = (dt2, launches2, ms2, byts2)
    finally:
        _native.set_option("max_ops_per_pass", -1)
        _native.set_option("jit", jit)
        _native.set_option(key, -1)
    k_best = min(per_k, key=lambda k: per_k[k][2])
    dt2, launches2, ms2, byts2 = per_k[k_best]
    gbs2 = byts2 / (ms2 * 1e-3) / 1e9
    res["unfused"] = {"gates": len(short), "seconds": round(dt2, 4), "hbm_gbs": round(gbs2, 1),
                      "frac_of_peak": round(gbs2 / HBM_PEAK_GBS, 4), "launches": launches2, "tile_qubits": k_best,
                      "hbm_gbs_by_tile_qubits": {k: round(v[3] / (v[2] * 1e-3) / 1e9, 1) for k, v in per_k.items()}}
    del sim2
    res["fusion"] = run_fusion_rows(dtype, n, circ, sim)
    return res, sim


def run_fusion_rows(dtype, n, circ, ref_sim):
    """Dense k-qubit fusion (k = 2..5) of the same circuit: blocks built on the host (untimed), then
    (a) one streaming dense sweep per block and (b) for k <= 4 the blocks as dense gates in the tile
    passes.  Each row: blocks, seconds, HBM GB/s of its kernels, and the fidelity defect against the
    unfused product-path state of the same precision; complex64 k = 5 also on the VALU instead of the
    matrix cores."""
    import mindquantum_amd as mq
    from mindquantum_amd import _native, fusion

    rows = {}
    state = (1 << n) * (8 if dtype == mq.complex64 else 16)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=1)

    def defect():
        return abs(1.0 - abs(mq.inner_product(ref_sim, sim)))

    def timed_apply(fn, kernels, reps=2):
        # the best of `reps` applications from |0> (the first after an untimed compile can find the
        # GPU's clocks still low); kernel figures per application
        times = []
        _native.profile_reset()
        _native.set_option("profile", 1)
        for _ in range(reps):
            sim.reset()
            t0 = time.perf_counter()
            fn()
            times.append(time.perf_counter() - t0)
        _native.set_option("profile", 0)
        dt = min(times)
        launches = ms = byts = 0
        for kname in kernels:
            a, b, c = _native.profile(kname)
            launches, ms, byts = launches + a // reps, ms + b / reps, byts + c / reps
        return {"seconds": round(dt, 4), "launches": launches, "kernel_ms": round(ms, 3),
                "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1) if ms else None,
                "frac_of_peak": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4) if ms else None,
                "gate_equivalent_gbs": round(len(circ) * 2 * state / dt / 1e9, 1)}

    for k in (2, 3, 4, 5):
        t0 = time.perf_counter()
        fc = fusion.fuse(circ, k)
        row = {"blocks": len(fc), "block_widths": fc.histogram(), "fuse_seconds_untimed": round(time.perf_counter() - t0, 2)}
        row["sweep"] = timed_apply(lambda: fc.apply(sim), ("dense_sweep", "dense_sweep_mfma"))
        row["sweep"]["fidelity_defect_vs_unfused"] = defect()
        if k == 5 and dtype == mq.complex64:
            _native.set_option("dense_mfma", 0)
            try:
                row["sweep_valu_only"] = timed_apply(lambda: fc.apply(sim), ("dense_sweep",))
                row["sweep_valu_only"]["fidelity_defect_vs_unfused"] = defect()
            finally:
                _native.set_option("dense_mfma", -1)
            _native.profile_reset()
        if k <= 4:
            tc = fc.as_circuit()
            sim.apply_circuit(tc)  # plan and compiled passes, untimed
            row["tile_passes"] = timed_apply(lambda: sim.apply_circuit(tc), ("tile_pass_fwd",))
            row["tile_passes"]["fidelity_defect_vs_unfused"] = defect()
        rows[f"k{k}"] = row
        log(f"fusion {n}q {np.dtype(dtype).name} k={k}: {row}")
    del sim
    return rows


def gate_apply_cpu(dtype, n, layers, depth, seed):
    """The CPU oracle on the first `layers` layers of the 28-qubit circuit (every layer costs the
    same: n one-qubit gates and n / 2 two-qubit gates, one sweep each), scaled to the full depth."""
    import mindquantum_amd as mq
    from mindquantum_amd import ir
    from mindquantum_amd.algorithm import layered_random_circuit
    from oracle import cpu

    circ = layered_random_circuit(n, layers, seed)
    cores = host_cores()
    cpu.set_num_threads(cores)
    qs = cpu.zero_state(n, dtype)
    t0 = time.perf_counter()
    cpu.apply_circuit(qs, ir.lower(circ, []), np.zeros(0))
    dt = time.perf_counter() - t0
    del qs
    return {"seconds_sample": round(dt, 2), "sample": f"{layers} of {depth} layers ({len(circ)} gates)",
            "seconds_full_depth_extrapolated": round(dt * depth / layers, 1), "cores": cores, "kind": "port"}


# ---------------------------------------------------------------- uneven QAOA batch
def run_uneven_items(count, seed, cpu_sample):
    """Uneven MaxCut batch: `count` QAOA circuits (p = 4) on 12..24 qubits, expectation and
    adjoint gradient of every circuit in complex64 through one mqa_grad_items call (shared
    per-step launches over the tiles of the smaller circuits, the largest on their own compiled
    passes; the compiles run in the untimed warm-up call).  Optionally the CPU oracle on a subset
    (two circuits of every size from 12 to 20 qubits) beside the GPU on the same subset, and the
    subset's error against the complex128 oracle."""
    import mindquantum_amd as mq
    from mindquantum_amd import ir
    from mindquantum_amd.algorithm import uneven_qaoa_batch

    circs, ops, nqs, params = uneven_qaoa_batch(count, seed)
    hams = [mq.Hamiltonian(o, mq.complex64) for o in ops]
    # algorithmic bytes of the reference schedule per circuit: (90.5 n + 8) 2^n 8 B
    ref_bytes = sum((90.5 * n + 8) * (1 << n) * 8 for n in nqs)
    mq.expectation_with_grad_batch(circs, hams, params, nqs, mq.complex64)  # plans, code objects
    log(f"uneven batch: warm-up done ({count} circuits)")
    t0 = time.perf_counter()
    f, grads = mq.expectation_with_grad_batch(circs, hams, params, nqs, mq.complex64)
    dt = time.perf_counter() - t0
    res = {"circuits": count, "qubits": [min(nqs), max(nqs)], "layers": 4, "seconds": round(dt, 4),
           "circuits_per_s": round(count / dt, 1),
           "reference_schedule_tb": round(ref_bytes / 1e12, 3),
           "reference_schedule_equivalent_gbs": round(ref_bytes / dt / 1e9, 1)}
    # roofline from this build's own schedule: a profiled repeat of the call (HIP events around every
    # shared launch, with the bytes its passes move: each item's state read and written once per
    # pass, psi and lambda in back-sweep passes)
    from mindquantum_amd import _native

    _native.profile_reset()
    _native.set_option("profile", 1)
    try:
        mq.expectation_with_grad_batch(circs, hams, params, nqs, mq.complex64)
    finally:
        _native.set_option("profile", 0)
    kern = {}
    # shared launches (items_*) and the large items' own compiled launches (tile_pass_*, hamiltonian)
    for name in ("items_tile_pass_fwd", "items_tile_pass_adj", "items_hamiltonian", "tile_pass_fwd", "tile_pass_adj",
                 "hamiltonian"):
        launches, ms, byts = _native.profile(name)
        if launches == 0:
            continue
        kern[name] = {"launches": launches, "ms": round(ms, 3), "gb": round(byts / 1e9, 2),
                      "gbs": round(byts / (ms * 1e-3) / 1e9, 1) if ms else None}
    tot_ms = sum(k["ms"] for k in kern.values())
    tot_gb = sum(k["gb"] for k in kern.values())
    res["roofline"] = {"bound": "hbm", "basis": "own schedule bytes / kernel time", "kernels": kern,
                       "kernel_ms": round(tot_ms, 3), "achieved_gbs": round(tot_gb / (tot_ms * 1e-3), 1),
                       "peak_gbs": HBM_PEAK_GBS, "frac": round(tot_gb / (tot_ms * 1e-3) / HBM_PEAK_GBS, 4)}
    log(f"uneven batch: {res}")
    if cpu_sample:
        from oracle import cpu

        idx = []  # two circuits of every size from 12 to 20 qubits
        for n in range(12, 21):
            idx += [i for i, m in enumerate(nqs) if m == n][:2]
        sub = ([circs[i] for i in idx], [hams[i] for i in idx], [params[i] for i in idx], [nqs[i] for i in idx])
        mq.expectation_with_grad_batch(*sub, mq.complex64)
        t0 = time.perf_counter()
        f_sub, g_sub = mq.expectation_with_grad_batch(*sub, mq.complex64)
        gpu_dt = time.perf_counter() - t0
        cores = host_cores()
        cpu.set_num_threads(cores)
        t0 = time.perf_counter()
        for i in idx:
            names = list(circs[i].params_name)
            cpu.grad_batch(nqs[i], np.complex64, ir.lower(circs[i], names), [hams[i].terms_ir()],
                           np.zeros((1, 0)), params[i], batch_threads=1)
        cpu_dt = time.perf_counter() - t0
        f_err, g_err, g_rel = 0.0, 0.0, 0.0
        for k, i in enumerate(idx):
            names = list(circs[i].params_name)
            fr, gr = cpu.grad_batch(nqs[i], np.complex128, ir.lower(circs[i], names),
                                    [mq.Hamiltonian(ops[i]).terms_ir()], np.zeros((1, 0)), params[i], batch_threads=1)
            f_err = max(f_err, abs(f_sub[k] - fr[0, 0]) / max(1.0, coeff_sum(ops[i])))
            dg = float(np.max(np.abs(g_sub[k] - gr[0, 0].real)))
            g_err = max(g_err, dg / max(1.0, coeff_sum(ops[i])))
            g_rel = max(g_rel, dg / float(np.max(np.abs(gr[0, 0].real))))
        sub_bytes = sum((90.5 * nqs[i] + 8) * (1 << nqs[i]) * 8 for i in idx)
        res["cpu_subset"] = {"circuits": len(idx), "qubits": [min(nqs[i] for i in idx), max(nqs[i] for i in idx)],
                             "cpu_seconds": round(cpu_dt, 3), "gpu_seconds": round(gpu_dt, 4), "cores": cores,
                             "kind": "port", "speedup": round(cpu_dt / gpu_dt, 1),
                             "cpu_seconds_full_batch_extrapolated": round(cpu_dt * ref_bytes / sub_bytes, 1)}
        res["err_vs_c128_oracle"] = {"items": len(idx), "f_rel_sum_c": f_err, "g_abs_over_sum_c": g_err,
                                     "g_rel_max_g": g_rel}
        log(f"uneven batch cpu subset: {res['cpu_subset']} err {res['err_vs_c128_oracle']}")
    return res


# ---------------------------------------------------------------- large single instances
def run_large_adjoint(dtype, n, layers, n_terms, seed, checkpoints):
    """One large instance: HEA of `layers` layers on n qubits (RY, RZ, CNOT ladder), 64 random
    Pauli strings of weight 2-4 with N(0,1) coefficients; one (f, grad) through the Simulator API."""
    import mindquantum_amd as mq
    from mindquantum_amd import _native
    from mindquantum_amd.algorithm import large_hea

    circ, op, ans = large_hea(n, layers, n_terms, seed)
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
    grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)
    _native.set_option("checkpoints", checkpoints)
    jit = _native.get_option("jit")
    _native.set_option("jit", 2)  # the warm-up waits for every compiled pass (checkpoint segments too)
    try:
        grad_ops(ans)  # plans and compiled passes
        _native.profile_reset()
        _native.set_option("profile", 1)
        t0 = time.perf_counter()
        f, g = grad_ops(ans)
        dt = time.perf_counter() - t0
        _native.set_option("profile", 0)
    finally:
        _native.set_option("checkpoints", -1)
        _native.set_option("jit", jit)
    amp = 8 if dtype == mq.complex64 else 16
    kern = {k: _native.profile(k) for k in ("tile_pass_fwd", "tile_pass_adj", "hamiltonian")}
    res = {"qubits": n, "dtype": np.dtype(dtype).name, "params": len(circ.params_name), "gates": len(circ),
           "terms": n_terms, "state_gb": round((1 << n) * amp / 1e9, 2), "seconds": round(dt, 3),
           "checkpoints": checkpoints, "f": float(np.real(f[0, 0])),
           "kernel_ms": {k: round(v[1], 1) for k, v in kern.items()},
           "kernel_hbm_gbs": {k: round(v[2] / (v[1] * 1e-3) / 1e9, 1) for k, v in kern.items() if v[1] > 0},
           "err": "no oracle at this size: see analog_24q (same circuit family) and the closed-form 32q / 31q "
                  "test in tests/test_gpu_fullsize.py"}
    log(f"large adjoint: {res}")
    del sim, grad_ops
    return res


def run_analog_24q(seed, cpu):
    """The large-instance circuit family at 24 qubits (HEA-4, 64 Pauli strings): GPU complex64 and
    complex128 beside the CPU oracle (b = 1, OpenMP over all cores), errors against complex128."""
    import mindquantum_amd as mq
    from mindquantum_amd import _native, ir
    from mindquantum_amd.algorithm import large_hea
    from oracle import cpu as ocpu

    n = 24
    circ, op, ans = large_hea(n, 4, 64, seed)
    names = list(circ.params_name)
    res = {"qubits": n, "layers": 4, "terms": 64, "params": len(names)}
    out = {}
    jit = _native.get_option("jit")
    _native.set_option("jit", 2)
    try:
        for dtype in (mq.complex64, mq.complex128):
            sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
            grad_ops = sim.get_expectation_with_grad(mq.Hamiltonian(op, dtype), circ)
            grad_ops(ans)
            t0 = time.perf_counter()
            f, g = grad_ops(ans)
            out[dtype] = (time.perf_counter() - t0, f[0, 0], g[0, 0].real)
            del sim, grad_ops
    finally:
        _native.set_option("jit", jit)
    res["gpu_seconds_c64"] = round(out[mq.complex64][0], 4)
    res["gpu_seconds_c128"] = round(out[mq.complex128][0], 4)
    cores = host_cores()
    ocpu.set_num_threads(cores)
    arrays = ir.lower(circ, names)
    terms = [mq.Hamiltonian(op).terms_ir()]
    t0 = time.perf_counter()
    f_ref, g_ref = ocpu.grad_batch(n, np.complex128, arrays, terms, np.zeros((1, 0)), ans, batch_threads=1)
    res["cpu_seconds_c128"] = round(time.perf_counter() - t0, 2)
    if cpu:
        t0 = time.perf_counter()
        ocpu.grad_batch(n, np.complex64, arrays, terms, np.zeros((1, 0)), ans, batch_threads=1)
        res["cpu_seconds_c64"] = round(time.perf_counter() - t0, 2)
        res["speedup_c64"] = round(res["cpu_seconds_c64"] / res["gpu_seconds_c64"], 1)
    res["speedup_c128"] = round(res["cpu_seconds_c128"] / res["gpu_seconds_c128"], 1)
    res["cpu_cores"] = cores
    csum = coeff_sum(op)
    for dtype, key in ((mq.complex64, "err_vs_c128_oracle_c64"), (mq.complex128, "err_vs_c128_oracle_c128")):
        _, f, g = out[dtype]
        res[key] = err_stats(f, g, f_ref[0, 0], g_ref[0, 0].real, csum)
    log(f"24q analog: {res}")
    return res


def copy_probe(gib=4, reps=10):
    """Plain device-to-device copy of `gib` GiB (the library's element-wise copy kernel, timed with
    HIP events): the achievable streaming rate on this GPU, and the kernel of known traffic that
    calibrates the HBM byte counters of the profiling runs."""
    import ctypes

    from mindquantum_amd import _native

    ms = ctypes.c_double(0)
    _native.check(_native.lib().mqa_copy_probe(gib << 30, reps, 0, ctypes.byref(ms)))
    byts = 2 * (gib << 30)
    res = {"bytes_per_copy": byts, "avg_ms": round(ms.value, 4), "gbs": round(byts / (ms.value * 1e-3) / 1e9, 1)}
    _native.check(_native.lib().mqa_copy_probe(gib << 30, reps, 1, ctypes.byref(ms)))
    res["in_place_gbs"] = round(byts / (ms.value * 1e-3) / 1e9, 1)
    _native.check(_native.lib().mqa_copy_probe(gib << 30, reps, 2, ctypes.byref(ms)))
    res["in_place_nontemporal_gbs"] = round(byts / (ms.value * 1e-3) / 1e9, 1)
    log(f"copy probe: {res}")
    return res


def cpu_baseline(n, layers, seed, samples):
    """The CPU oracle (reference mqvector algorithm restated, C++/OpenMP, -O3 -mavx2) on `samples`
    rows of the headline workload in the three thread layouts of the reference (ParallelTask
    workers x OpenMP threads per worker, gradient.h:270-299): one worker per core with one OpenMP
    thread each; one worker with OpenMP over all cores; the reference default (parallel_worker=None:
    one worker per sample, each with the OpenMP default of all cores).  The best is reported."""
    import mindquantum_amd as mq
    from mindquantum_amd import ir
    from mindquantum_amd.algorithm import batched_adjoint
    from oracle import cpu

    circ, op, enc, ans = batched_adjoint(n, layers, samples, seed)
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    arrays = ir.lower(circ, names)
    terms = [mq.Hamiltonian(op).terms_ir()]
    cores = host_cores()
    cpu.set_num_threads(cores)
    layouts = {"thread_per_sample": (cores, 1), "omp_per_sample": (1, cores), "reference_default": (samples, cores)}
    rates, secs = {}, {}
    for name, (workers, omp) in layouts.items():
        log(f"cpu baseline: {samples} samples, {workers} workers x {omp} OpenMP threads ({name})")
        t0 = time.perf_counter()
        cpu.grad_batch(n, np.complex64, arrays, terms, enc, ans, batch_threads=workers, worker_threads=omp)
        secs[name] = time.perf_counter() - t0
        rates[name] = samples / secs[name]
    best = max(rates, key=rates.get)
    log(f"cpu baseline: {rates} ({cores} cores)")
    return {
        "value": round(rates[best], 4), "unit": "grads/s", "cores": cores, "kind": "port",
        "sample": f"{n}q c64 HEA-{layers} adjoint gradients, {samples} samples per layout: "
                  + "; ".join(f"{k} {layouts[k][0]} workers x {layouts[k][1]} OpenMP ({secs[k]:.1f} s, "
                              f"{rates[k]:.3f}/s)" for k in layouts) + f"; best = {best}",
        "layouts": {k: round(v, 4) for k, v in rates.items()},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--batch", type=int, default=1024)
    ap.add_argument("--qubits", type=int, default=20)
    ap.add_argument("--layers", type=int, default=8)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true")
    ap.add_argument("--apply-qubits", type=int, default=28)
    ap.add_argument("--apply-depth", type=int, default=200)
    ap.add_argument("--interpreted", action="store_true", help="run the interpreted pass kernels (no code generation)")
    ap.add_argument("--items", type=int, default=512, help="circuits of the uneven QAOA batch (0: skip)")
    ap.add_argument("--no-large", action="store_true", help="skip the 32q c64 / 31q c128 single-instance gradients")
    ap.add_argument("--large-checkpoints", type=int, default=4)
    ap.add_argument("--c128-steps", type=int, default=5)
    ap.add_argument("--cpu-samples", type=int, default=64)
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile widths (no autotuning)")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the device copy probe (counter calibration for profiling runs)")
    ap.add_argument("--apply-only", action="store_true",
                    help="run only the complex64 gate application of the random circuit, fused passes (profiling)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="set a library option (mqa_set_option) before the workloads; repeatable")
    ap.add_argument("--mark-workloads", action="store_true",
                    help="log each workload's start time (host clocks), to split a kernel trace by workload")
    args = ap.parse_args()
    global MARKS
    MARKS = args.mark_workloads

    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build runs every workload on a single MI355X "
              "(multi-GPU sharding is a design note, not a measured path). Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import mindquantum_amd as mq
    from mindquantum_amd import _native

    if _native.device_count() < 1:
        print("no GPU visible", file=sys.stderr)
        sys.exit(1)
    if args.copy_probe_only:
        print(json.dumps({"copy_probe": copy_probe()}), flush=True)
        return
    if args.apply_only:
        _native.set_option("jit", 2)
        for kv in args.option:
            key, _, val = kv.partition("=")
            _native.set_option(key.strip(), int(val))
        res, _ = run_gate_apply(mq.complex64, args.apply_qubits, args.apply_depth, args.seed, 0, variants=False)
        print(json.dumps({"gate_apply_c64": res}), flush=True)
        return
    # per-circuit compiled passes, compiled before the first launch (inside the warmup steps)
    _native.set_option("jit", 0 if args.interpreted else 2)
    for kv in args.option:
        key, _, val = kv.partition("=")
        _native.set_option(key.strip(), int(val))

    tune = not args.no_tune and not args.interpreted
    mark("adjoint_20q_b1024_c64")
    h64 = run_adjoint(mq.complex64, args.qubits, args.layers, args.batch, args.steps, args.warmup, args.seed,
                      profile=True, tune=tune)
    value, ms_step, stats = h64["grads_per_s"], h64["ms_per_step"], h64["stats"]
    log(f"c64: {value:.1f} grads/s, {ms_step:.1f} ms/step")
    launches, ms, byts = stats["tile_pass_adj"]
    avg_ms = ms / max(launches, 1)
    per_launch = byts / max(launches, 1)
    achieved = per_launch / (avg_ms * 1e-3) / 1e9
    roofline = {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic() if not args.interpreted and (args.qubits, args.batch, args.layers) ==
                (20, 1024, 8) else None,
                "kernel": "k_tile_pass<float, adjoint>" if args.interpreted else "mqa_jit_tile_pass_adj (complex64)",
                "avg_launch_ms": round(avg_ms, 4),
                "algorithmic_bytes_per_launch": per_launch, "launches": launches}
    share = {k: round(v[1], 2) for k, v in stats.items()}
    extra = {"kernel_ms_in_timed_region": share, "compile_seconds_untimed": _native.get_option("jit_compile_ms") / 1e3,
             "autotune": h64["tiles"]}
    # the other two kernel families of the step, on the same basis (algorithmic bytes / launch time)
    for name in ("tile_pass_fwd", "hamiltonian"):
        if name in stats and stats[name][0] > 0:
            n_l, ms_l, b_l = stats[name]
            gbs = b_l / n_l / (ms_l / n_l * 1e-3) / 1e9
            extra[f"roofline_{name}"] = {"launches": n_l, "avg_launch_ms": round(ms_l / n_l, 4),
                                         "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}
    samples = [h64["sample"]]
    if not args.no_extra:
        mark("copy_probe_4GiB")
        extra["copy_probe"] = copy_probe()
        roofline["achievable_copy_gbs"] = extra["copy_probe"]["gbs"]
        mark("adjoint_20q_b1024_c128")
        h128 = run_adjoint(mq.complex128, args.qubits, args.layers, args.batch, args.c128_steps, 1, args.seed,
                           profile=False, tune=tune)
        samples.append(h128["sample"])
        extra["c128"] = {"grads_per_s": round(h128["grads_per_s"], 2), "ms_per_step": round(h128["ms_per_step"], 2),
                         "steps": args.c128_steps, "autotune": h128["tiles"]}
        log(f"c128: {h128['grads_per_s']:.1f} grads/s")
    if args.batch >= 16:
        errs = headline_errors(samples, args.qubits)
        extra["err_vs_c128_oracle"] = errs[0]
        if len(errs) > 1:
            extra["c128"]["err_vs_c128_oracle"] = errs[1]
    if not args.no_extra:
        mark("hea16_c128_single")
        extra["hea16_c128_single"] = run_hea16(20, not args.no_cpu)
        mark("gate_apply_28q_c64")
        g64, s64 = run_gate_apply(mq.complex64, args.apply_qubits, args.apply_depth, args.seed, 10, tune=tune)
        mark("gate_apply_28q_c128")
        g128, s128 = run_gate_apply(mq.complex128, args.apply_qubits, args.apply_depth, args.seed, 10, tune=tune)
        up = s64.astype(mq.complex128)
        g64["fidelity_defect_vs_c128_run"] = abs(1.0 - abs(mq.inner_product(s128, up)))
        del up, s64, s128
        if not args.no_cpu:
            g64["cpu_baseline"] = gate_apply_cpu(np.complex64, args.apply_qubits, 2, args.apply_depth, args.seed)
            g128["cpu_baseline"] = gate_apply_cpu(np.complex128, args.apply_qubits, 2, args.apply_depth, args.seed)
            for g in (g64, g128):
                g["speedup_vs_cpu"] = round(g["cpu_baseline"]["seconds_full_depth_extrapolated"] / g["seconds"], 1)
            log(f"28q cpu: {g64['cpu_baseline']} / {g128['cpu_baseline']}")
        extra["gate_apply_28q_c64"] = g64
        extra["gate_apply_28q_c128"] = g128
        if args.items > 0:
            mark("uneven_qaoa_512_c64")
            extra["uneven_qaoa_c64"] = run_uneven_items(args.items, args.seed, not args.no_cpu)
    if not args.no_large and not args.no_extra:
        mark("large_32q_c64")
        extra["large_32q_c64"] = run_large_adjoint(mq.complex64, 32, 4, 64, args.seed, args.large_checkpoints)
        mark("large_31q_c128")
        extra["large_31q_c128"] = run_large_adjoint(mq.complex128, 31, 4, 64, args.seed, args.large_checkpoints)
        mark("analog_24q")
        extra["analog_24q"] = run_analog_24q(args.seed, not args.no_cpu)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.qubits, args.layers, args.seed, args.cpu_samples)
        extra["speedup_vs_cpu"] = round(value / cpu["value"], 1)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "grads/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"{args.qubits}q encoder-RX + HEA-{args.layers} adjoint gradient, batch {args.batch}, "
                               f"complex64 (40-term ZZ ring + X Hamiltonian)",
                   "qubits": args.qubits, "batch": args.batch, "params": h64["n_params"], "gates": h64["n_gates"]},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
