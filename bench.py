#!/usr/bin/env python3
"""Benchmark of the MI355X state-vector backend.

Headline (``value``): batched adjoint gradients per second - 20 qubits, batch 1024, complex64:
encoder RX(x_i) on every qubit (x ~ N(0,1)), 8-layer hardware-efficient ansatz (RY, RZ on every
qubit, CNOT ladder i -> i+1; 320 parameters ~ U[0, 2pi)), H = sum_i w_i Z_i Z_{i+1 mod 20} +
sum_i v_i X_i with w, v ~ U(-1, 1) (40 terms).  One step = one forward + backward of the
torch.autograd operator (MQOps) over the whole batch; the adjoint sweep runs in the forward.

Also reported in the same JSON line: the same workload in complex128; 28-qubit gate
application (layered random circuit, depth 200, complex64 / complex128) in HBM GB/s, fused
(tile passes) and unfused (one gate per pass); the uneven batch of 512 MaxCut QAOA circuits
(12..24 qubits, p = 4, complex64, one mqa_grad_items call); one 32-qubit complex64 and one
31-qubit complex128 gradient (HEA-4, 64 random Pauli strings, 4 checkpoints).  ``roofline`` is the dominant kernel of the
headline (the adjoint tile pass) timed with HIP events on the library's stream;
``cpu_baseline`` times the CPU oracle (a restatement of the reference CPU mqvector path) on a
bounded sample of the same workload on this host's cores.
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


def measured_traffic():
    try:
        with open(TRAFFIC_FILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_launch"):
                    return float(line.split("=")[1])
    except OSError:
        pass
    return None
METRIC = "adjoint grads/s (20q, batch 1024, c64 vs c128); 28q gate-apply HBM GB/s"


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def adjoint_workload(n, layers, batch, seed):
    import mindquantum_amd as mq
    from mindquantum_amd.algorithm import hardware_efficient_ansatz
    from mindquantum_amd.algorithm.builders import encoder_rx

    rng = np.random.default_rng(seed)
    circ = encoder_rx(n) + hardware_efficient_ansatz(n, layers)
    op = mq.QubitOperator("", 0)
    w = rng.uniform(-1, 1, size=n)
    v = rng.uniform(-1, 1, size=n)
    for i in range(n):
        op += mq.QubitOperator(f"Z{i} Z{(i + 1) % n}", float(w[i]))
        op += mq.QubitOperator(f"X{i}", float(v[i]))
    enc = rng.normal(size=(batch, n))
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.ansatz_params_name))
    return circ, op, enc, ans


def run_adjoint(dtype, n, layers, batch, steps, warmup, seed, profile, tune=False):
    import torch

    import mindquantum_amd as mq
    from mindquantum_amd import _native, autotune
    from mindquantum_amd.framework import MQOps

    circ, op, enc, ans = adjoint_workload(n, layers, batch, seed)
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
    n_params = len(circ.params_name)
    if tune:
        stats["tiles"] = tiles
    return batch * steps / dt, dt / steps * 1e3, stats, n_params, len(circ)


def run_gate_apply(dtype, n, depth, seed, unfused_depth, tune=False):
    import mindquantum_amd as mq
    from mindquantum_amd import _native, autotune, ir
    from mindquantum_amd.algorithm import layered_random_circuit

    amp = 8 if dtype == mq.complex64 else 16
    circ = layered_random_circuit(n, depth, seed)
    tuned = None
    if tune:  # untimed: the tile width of the passes, chosen on this circuit
        times = autotune.tune_apply(circ, n, dtype)
        tuned = {"tile_qubits": circ._mqa_tiles[(mq.dtype.native_code(dtype), False)], "seconds": {k: round(v, 4) for k, v in times.items()}}
        log(f"autotune apply {n}q {np.dtype(dtype).name}: {tuned}")
    sim = mq.Simulator("mqvector_amd", n, dtype=dtype, seed=seed)
    code = 0 if dtype == mq.complex64 else 1
    passes = len(ir.NativeCircuit(circ, []).schedule(n, code))
    # untimed run of the same circuit: plan, allocations and the compiled passes (jit = 2)
    c0 = _native.get_option("jit_compile_ms")
    sim.apply_circuit(circ)
    compile_s = (_native.get_option("jit_compile_ms") - c0) / 1e3
    sim.reset()
    _native.profile_reset()
    _native.set_option("profile", 1)
    t0 = time.perf_counter()
    sim.apply_circuit(circ)
    dt = time.perf_counter() - t0
    _native.set_option("profile", 0)
    launches, ms, byts = _native.profile("tile_pass_fwd")
    state = (1 << n) * amp
    res = {
        "gates": len(circ), "passes": passes, "seconds": round(dt, 4),
        "hbm_gbs": round(byts / (ms * 1e-3) / 1e9, 1),
        "frac_of_peak": round(byts / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
        "gate_equivalent_gbs": round(len(circ) * 2 * state / dt / 1e9, 1),
        "kernel_ms": round(ms, 3), "launches": launches,
        "compiled_passes": _native.get_option("jit") != 0, "compile_seconds_untimed": round(compile_s, 1),
        "autotune": tuned,
    }
    log(f"gate apply {n}q {np.dtype(dtype).name}: {res}")
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
    del sim, sim2
    return res


def run_uneven_items(count, seed, cpu_sample):
    """Uneven MaxCut batch: `count` QAOA circuits (p = 4) on 12..24 qubits, expectation and
    adjoint gradient of every circuit in complex64 through one mqa_grad_items call (shared
    per-step launches over the tiles of all circuits).  Optionally the CPU oracle on a subset
    (circuits of at most 18 qubits) beside the GPU on the same subset."""
    import mindquantum_amd as mq
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
    log(f"uneven batch: {res}")
    if cpu_sample:
        from mindquantum_amd import ir
        from oracle import cpu

        idx = []  # two circuits of every size from 12 to 20 qubits
        for n in range(12, 21):
            idx += [i for i, m in enumerate(nqs) if m == n][:2]
        sub = ([circs[i] for i in idx], [hams[i] for i in idx], [params[i] for i in idx], [nqs[i] for i in idx])
        mq.expectation_with_grad_batch(*sub, mq.complex64)
        t0 = time.perf_counter()
        mq.expectation_with_grad_batch(*sub, mq.complex64)
        gpu_dt = time.perf_counter() - t0
        cores = len(os.sched_getaffinity(0))
        cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
        cpu.set_num_threads(cores)
        t0 = time.perf_counter()
        for i in idx:
            names = list(circs[i].params_name)
            cpu.grad_batch(nqs[i], np.complex64, ir.lower(circs[i], names), [hams[i].terms_ir()],
                           np.zeros((1, 0)), params[i], batch_threads=1)
        cpu_dt = time.perf_counter() - t0
        sub_bytes = sum((90.5 * nqs[i] + 8) * (1 << nqs[i]) * 8 for i in idx)
        res["cpu_subset"] = {"circuits": len(idx), "qubits": [min(nqs[i] for i in idx), max(nqs[i] for i in idx)],
                             "cpu_seconds": round(cpu_dt, 3), "gpu_seconds": round(gpu_dt, 4), "cores": cores,
                             "kind": "port", "speedup": round(cpu_dt / gpu_dt, 1),
                             "cpu_seconds_full_batch_extrapolated": round(cpu_dt * ref_bytes / sub_bytes, 1)}
        log(f"uneven batch cpu subset: {res['cpu_subset']}")
    return res


def run_large_adjoint(dtype, n, layers, n_terms, seed, checkpoints):
    """One large instance: HEA of `layers` layers on n qubits (RY, RZ, CNOT ladder), 64 random
    Pauli strings of weight 2-4 with N(0,1) coefficients; one (f, grad) through the Simulator API."""
    import mindquantum_amd as mq
    from mindquantum_amd import _native
    from mindquantum_amd.algorithm import hardware_efficient_ansatz, random_pauli_hamiltonian

    rng = np.random.default_rng(seed)
    circ = hardware_efficient_ansatz(n, layers)
    op = random_pauli_hamiltonian(n, n_terms, seed)
    ans = rng.uniform(0, 2 * np.pi, size=len(circ.params_name))
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
           "kernel_hbm_gbs": {k: round(v[2] / (v[1] * 1e-3) / 1e9, 1) for k, v in kern.items() if v[1] > 0}}
    log(f"large adjoint: {res}")
    del sim, grad_ops
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


def cpu_baseline(n, layers, seed, budget_s):
    """The CPU oracle (reference mqvector algorithm restated, C++/OpenMP, -O3 -mavx2) on a
    bounded sample of the headline workload; best of two thread layouts."""
    import mindquantum_amd as mq
    from mindquantum_amd import ir
    from oracle import cpu

    circ, op, enc, ans = adjoint_workload(n, layers, 64, seed)
    names = list(circ.encoder_params_name) + list(circ.ansatz_params_name)
    arrays = ir.lower(circ, names)
    terms = [mq.Hamiltonian(op).terms_ir()]
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    results = {}
    # (ii) batch_threads = 1, OpenMP over all cores inside each sample
    log(f"cpu baseline: 1 sample, OpenMP over {cores} threads")
    cpu.set_num_threads(cores)
    t0 = time.perf_counter()
    cpu.grad_batch(n, np.complex64, arrays, terms, enc[:1], ans, batch_threads=1)
    t1 = time.perf_counter() - t0
    results["omp_per_sample"] = 1 / t1
    # (i) batch_threads = cores, one OpenMP thread per sample
    log(f"cpu baseline: {cores} samples, one thread each")
    cpu.set_num_threads(1)
    m = max(cores, 1)
    t0 = time.perf_counter()
    cpu.grad_batch(n, np.complex64, arrays, terms, enc[:m], ans, batch_threads=m)
    t2 = time.perf_counter() - t0
    results["thread_per_sample"] = m / t2
    cpu.set_num_threads(cores)
    best = max(results, key=results.get)
    log(f"cpu baseline: {results} ({cores} cores)")
    return {
        "value": round(results[best], 4), "unit": "grads/s", "cores": cores, "kind": "port",
        "sample": f"{n}q c64 HEA-{layers} adjoint gradients: 1 sample with OpenMP over {cores} threads "
                  f"({t1:.1f} s, {results['omp_per_sample']:.3f}/s) and {m} samples one thread each ({t2:.1f} s, "
                  f"{results['thread_per_sample']:.3f}/s); best = {best}",
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
    ap.add_argument("--no-tune", action="store_true", help="keep the default tile widths (no autotuning)")
    ap.add_argument("--copy-probe-only", action="store_true",
                    help="run only the device copy probe (counter calibration for profiling runs)")
    ap.add_argument("--option", action="append", default=[], metavar="KEY=VALUE",
                    help="set a library option (mqa_set_option) before the workloads; repeatable")
    args = ap.parse_args()

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
    # per-circuit compiled passes, compiled before the first launch (inside the warmup steps)
    _native.set_option("jit", 0 if args.interpreted else 2)
    for kv in args.option:
        key, _, val = kv.partition("=")
        _native.set_option(key.strip(), int(val))

    tune = not args.no_tune and not args.interpreted
    value, ms_step, stats, n_params, n_gates = run_adjoint(mq.complex64, args.qubits, args.layers, args.batch,
                                                           args.steps, args.warmup, args.seed, profile=True,
                                                           tune=tune)
    tiles = stats.pop("tiles", None)
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
             "autotune": tiles}
    if not args.no_extra:
        extra["copy_probe"] = copy_probe()
        roofline["achievable_copy_gbs"] = extra["copy_probe"]["gbs"]
        v128, ms128, st128, _, _ = run_adjoint(mq.complex128, args.qubits, args.layers, args.batch, 1, 1,
                                               args.seed, profile=False, tune=tune)
        extra["c128_autotune"] = st128.get("tiles")
        extra["c128_grads_per_s"] = round(v128, 2)
        extra["c128_ms_per_step"] = round(ms128, 2)
        log(f"c128: {v128:.1f} grads/s")
        extra["gate_apply_28q_c64"] = run_gate_apply(mq.complex64, args.apply_qubits, args.apply_depth, args.seed, 10,
                                                     tune=tune)
        extra["gate_apply_28q_c128"] = run_gate_apply(mq.complex128, args.apply_qubits, args.apply_depth, args.seed,
                                                      10, tune=tune)
        if args.items > 0:
            extra["uneven_qaoa_c64"] = run_uneven_items(args.items, args.seed, not args.no_cpu)
    if not args.no_large and not args.no_extra:
        extra["large_32q_c64"] = run_large_adjoint(mq.complex64, 32, 4, 64, args.seed, args.large_checkpoints)
        extra["large_31q_c128"] = run_large_adjoint(mq.complex128, 31, 4, 64, args.seed, args.large_checkpoints)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.qubits, args.layers, args.seed, 20)
        extra["speedup_vs_cpu"] = round(value / cpu["value"], 1)

    line = {
        "metric": METRIC, "value": round(value, 2), "unit": "grads/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms_step, 3), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "c64", "data": "synthetic",
        "config": {"workload": f"{args.qubits}q encoder-RX + HEA-{args.layers} adjoint gradient, batch {args.batch}, "
                               f"complex64 (40-term ZZ ring + X Hamiltonian)",
                   "qubits": args.qubits, "batch": args.batch, "params": n_params, "gates": n_gates},
        "roofline": roofline, "cpu_baseline": cpu, "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
