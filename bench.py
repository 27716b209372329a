"""Throughput of the state-vector engine on one MI355X.

Headline (the `value` of the JSON line): the 30-qubit complex-f32 random
circuit (depth 40; Rx, Ry, Rz on every qubit then 10 CNOT/CZ pairs per layer).
A step = one full simulation of the circuit from |0..0> through the fused
schedule (every gate applied; the state stays in HBM); `value` counts the
circuit's original gate applications per second.

The same line carries every workload of the scope table:
  roofline          the fused stage kernel: algorithmic bytes (2 * 2^30 * 8 B
                    per stage) / mean stage duration from HIP events in the
                    timed region, against the 8 TB/s HBM peak;
  roofline_valu     the same kernel's FP32 vector work (from the packed
                    program's records) per stage duration, against the
                    78.6 TFLOP/s FP32 vector peak without packed FP32: the
                    bound the stage actually sits on;
  per_gate_1q_GBps  the exact single-qubit gate kernel, GB/s for every target;
  f32_error_vs_f64  the complex-f32 fused state against the exact complex-f64
                    path (bit-identical to the FP64 oracle) at 30 qubits, and
                    (exact_c32) the unfused complex-f32 path, one exact kernel
                    per gate: its gate-apps/s and its error;
  dense_blocks      the circuit as dense 5-qubit blocks on the matrix cores;
  c64_32q           the same generator at 32 qubits complex f64, depth 20;
  cpu_baseline      the reference's fastest CPU path: simd.rs's AVX2 +
                    thread-pool kernels (complex f64, restated in the oracle)
                    on the leading gates of the headline circuit, with the host
                    threads of this process (at most 16); the f32 scalar path
                    (QuantumStateF32, one thread) beside it;
  config1           20 qubits complex f64 depth 20: amplitudes, <Z_i> for every
                    qubit and 10^4 shots on the device, and the CPU paths
                    (scalar and simd.rs; run_shots as written, extrapolated,
                    and as measure_all on copies);
  config3           QAOA MaxCut on a 26-node 3-regular graph, p = 4, complex
                    f64: <H_C> and 10^5 sampled cuts on the device, and the
                    CPU paths (sampled gates, extrapolated);
  config4           4096 QAOA (n ~ U{10..22}, G(n, 0.5), p ~ U{1..5}) + 512 VQE
                    (n ~ U{12..16}, Ry-CZ ansatz, 4 layers, 100 Pauli strings)
                    in one persistent batch: instances/s of one evaluation and
                    of one optimizer iteration (hipGraph replay), and the CPU
                    evaluation rate on a sample;
  qaoa              the learned warm start on held-out MaxCut instances
                    (G(n, 0.5), n ~ U{10..22}, U(0,1) weights, p ~ U{1..8})
                    against the plain batched optimizer (run_qaoa's rule, 16
                    iterations): instances/s of both and the fraction within
                    1 % of the plain energy.
CPU legs run the oracle's restatements of the reference's CPU paths on the
host cores of this machine; nothing else in this file uses the oracle.
"""
import argparse
import ctypes as C
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
VALU_FP32_PEAK_TFLOPS = 78.6  # FP32 vector FMA rate without packed FP32 (half of 157.3)
MFMA_F32_PEAK_TFLOPS = 157.3

# FP32 flops per amplitude of one packed record of the fused program, by kind:
# dense 2x2, dense 4x4, 1-bit diagonal, 2-bit diagonal, CNOT, CZ, padding,
# factored 2x2, pending-diagonal table (an FMA counts as 2).
FLOPS_PER_AMP = [16, 32, 6, 6, 0, 0, 0, 8, 6]


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_threads():
    return max(1, min(len(os.sched_getaffinity(0)), 16))


# ---- CPU legs: the oracle's restatements of the reference's CPU paths -------

def cpu_gates(nq, arr, budget_s, path, threads=1, f32=False):
    """Applies the leading gates of `arr` to |0..0> on one CPU path until
    `budget_s` of work is done: path 'scalar' (state.rs:214-276, or
    mixed_precision.rs:396-458 with f32) or 'simd' (simd.rs *_best).
    Returns (gates done, seconds)."""
    from oracle import oracle as orc
    amps = np.zeros(1 << nq, dtype=np.complex64 if f32 else np.complex128)
    amps[0] = 1
    lib = orc.lib()
    p = amps.ctypes.data_as(C.c_void_p)
    cur, nout = C.c_size_t(0), C.c_size_t(0)
    two = (orc.CNOT, orc.CZ, orc.SWAP, orc.RZZ)
    done = 0
    t0 = time.perf_counter()
    while done < len(arr) and time.perf_counter() - t0 < budget_s:
        g = np.ascontiguousarray(arr[done:done + 1]).view(orc.GATE_DTYPE)
        if path == "scalar":
            if lib.orc_apply_gate(p, 1 if f32 else 0, nq, g.ctypes.data_as(C.c_void_p), None, 0, C.byref(cur), None,
                                  C.byref(nout)):
                raise RuntimeError("oracle gate failed")
        elif int(g[0]["kind"]) in two:
            orc.simd_apply_2q_best(amps, nq, int(g[0]["q0"]), int(g[0]["q1"]), orc.matrix_2q(g[0]), threads)
        else:
            orc.simd_apply_1q_best(amps, nq, int(g[0]["q0"]), orc.matrix_1q(g[0]), threads)
        done += 1
    return done, time.perf_counter() - t0


def cpu_measure_all(nq, shots_budget_s, amps):
    """measure_all (state.rs:329-335) on fresh copies of `amps`, one shot at a
    time, until the budget is spent.  Returns seconds per shot."""
    from oracle import oracle as orc
    draws = np.random.default_rng(0).random((64, nq))
    t0 = time.perf_counter()
    k = 0
    while k < 64 and (k == 0 or time.perf_counter() - t0 < shots_budget_s):
        a = amps.copy()
        orc.measure_all(a, nq, draws[k])
        k += 1
    return (time.perf_counter() - t0) / k


def measured_value(path, key):
    """A figure from a committed profile summary (None when absent)."""
    try:
        for line in open(path):
            if line.startswith(key):
                return float(line.split("=")[1])
    except (OSError, ValueError):
        pass
    return None


# ---- the headline: 30 qubits complex f32, fused -------------------------------

def headline(args, L):
    from ruqu_amd import QuantumState
    from ruqu_amd._native import Metrics, ptr
    from ruqu_amd import workloads

    nq, prec = args.qubits, args.precision
    amp_bytes = 8 if prec == "c32" else 16
    log("building circuit %dq depth %d" % (nq, args.depth))
    circ = workloads.random_circuit(nq, args.depth, args.seed)
    arr = circ.to_array()
    ngates = len(arr)
    st = QuantumState.new_with_seed(nq, args.seed, prec)
    met = Metrics()
    nout = C.c_size_t(0)

    def step():
        rc = L.rq_run_circuit(st.handle, ptr(arr), ngates, None, 0, None, 0, C.byref(nout), 1, C.byref(met))
        if rc:
            raise RuntimeError(L.rq_last_error().decode())
        return met.device_time_ms, met.passes

    for w in range(args.warmup):
        step()
        log("warmup %d done" % w)
    L.rq_device_synchronize()
    dev_ms, passes = [], 0
    t0 = time.perf_counter()
    for k in range(args.steps):
        ms, passes = step()
        dev_ms.append(ms)
    L.rq_device_synchronize()
    el = time.perf_counter() - t0
    log("timed %d steps in %.3f s (%d passes/step)" % (args.steps, el, passes))

    n = 1 << nq
    stage_bytes = 2.0 * n * amp_bytes
    mean_stage_ms = float(np.mean(dev_ms)) / passes
    achieved = stage_bytes / (mean_stage_ms * 1e-3) / 1e9
    out = {"ngates": ngates, "el": el, "passes": passes, "arr": arr}
    out["roofline"] = {
        "bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": (measured_value(os.path.join(ROOT, "profiles", "fused_stage_30q_c32_hbm_traffic.md"),
                                   "traffic_bytes_per_launch") if (nq, prec, args.depth) == (30, "c32", 40) else None),
        "kernel": "k_fused_stage", "bytes_per_launch": stage_bytes, "mean_launch_ms": round(mean_stage_ms, 4),
        "limiter": "VALU issue and the scalar-load and LDS latency 4 waves per SIMD leave exposed (roofline_valu)"}

    # FP32 vector work of the packed program (the matrix-core phases apart)
    cnt = np.zeros(9, dtype=np.uint64)
    if prec == "c32" and L.rq_fusion_packed_kinds(nq, 1, ptr(arr), ngates, L.rq_tune_get(2), L.rq_tune_get(3),
                                                  ptr(cnt)) == 0:
        flops = float(sum(int(c) * f for c, f in zip(cnt, FLOPS_PER_AMP))) * n / passes
        tf = flops / (mean_stage_ms * 1e-3) / 1e12
        instr = measured_value(os.path.join(ROOT, "profiles", "fused_stage_30q_c32_counters.md"),
                               "valu_instr_per_wave_per_stage")
        out["roofline_valu"] = {
            "bound": "valu", "achieved": round(tf, 2), "peak": VALU_FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": round(tf / VALU_FP32_PEAK_TFLOPS, 4), "kernel": "k_fused_stage",
            "fp32_flops_per_launch": flops,
            "records": {k: int(c) for k, c in zip(["u1", "u2", "d1", "d2", "cx", "cz", "pad", "u1f", "dtab"], cnt)},
            "valu_instr_per_wave_per_stage": instr,
            "note": "FP32 arithmetic of the packed records per stage duration; the VALU also issues the register "
                    "moves, LDS and address instructions, so the issue fraction (profiles/) is higher"}

    # exact single-qubit kernel bandwidth, every target qubit
    per_gate = {}
    m = np.array([[0.6, 0.8j], [0.8j, 0.6]], dtype=np.complex128)
    for q in range(nq):
        avg = C.c_double()
        L.rq_time_1q(st.handle, q, ptr(m), 1, C.byref(avg))  # warm
        if L.rq_time_1q(st.handle, q, ptr(m), 5, C.byref(avg)):
            raise RuntimeError(L.rq_last_error().decode())
        per_gate[str(q)] = round(stage_bytes / (avg.value * 1e-3) / 1e9, 1)
    cavg = C.c_double()
    copy_gbs = None
    if L.rq_time_copy(st.handle, 5, C.byref(cavg)) == 0:
        copy_gbs = round(stage_bytes / (cavg.value * 1e-3) / 1e9, 1)
    vals = list(per_gate.values())
    log("per-gate GB/s min %.0f max %.0f, copy %s GB/s" % (min(vals), max(vals), copy_gbs))
    out["per_gate_1q_GBps"] = per_gate
    out["per_gate_1q_min_GBps"] = min(vals)
    out["copy_GBps"] = copy_gbs

    err = None
    if prec == "c32" and not args.no_f64_check and nq <= 31:
        log("exact f64 run of the same circuit for the f32 error ...")
        step()  # the per-gate timing above changed the state: run the circuit again
        a32 = st.state_vector()
        st64 = QuantumState.new_with_seed(nq, args.seed, "c64")
        t1 = time.perf_counter()
        rc = L.rq_run_circuit(st64.handle, ptr(arr), ngates, None, 0, None, 0, C.byref(nout), 0, C.byref(met))
        if rc:
            raise RuntimeError(L.rq_last_error().decode())
        exact_s = time.perf_counter() - t1
        a64 = st64.state_vector()
        del st64
        # the unfused complex-f32 path (the reference's QuantumStateF32 arithmetic, one kernel per gate)
        t2 = time.perf_counter()
        rc = L.rq_run_circuit(st.handle, ptr(arr), ngates, None, 0, None, 0, C.byref(nout), 0, C.byref(met))
        if rc:
            raise RuntimeError(L.rq_last_error().decode())
        exact32_s = time.perf_counter() - t2
        exact32_dev_ms = met.device_time_ms
        a32x = st.state_vector()

        def err_stats(a_32, a_64):
            mx, d2, n64, n32 = 0.0, 0.0, 0.0, 0.0
            ip = 0j
            ch = 1 << 24
            for o in range(0, n, ch):
                x = a_32[o:o + ch].astype(np.complex128)
                y = a_64[o:o + ch]
                d = x - y
                mx = max(mx, float(np.max(np.abs(d))))
                d2 += float(np.vdot(d, d).real)
                n64 += float(np.vdot(y, y).real)
                n32 += float(np.vdot(x, x).real)
                ip += complex(np.vdot(y, x))
            return {"max_abs": mx, "rel_l2": (d2 / n64) ** 0.5, "one_minus_fidelity": 1.0 - abs(ip) ** 2 / (n32 * n64)}

        err = err_stats(a32, a64)
        err["exact_f64_gate_apps_per_s"] = round(ngates / exact_s, 1)
        err["note"] = ("complex-f32 fused run vs the exact per-gate complex-f64 run of the same circuit on the "
                       "device (that path is bit-identical to the FP64 CPU oracle wherever the oracle runs)")
        e32 = err_stats(a32x, a64)
        e32.update(gate_apps_per_s=round(ngates / exact32_s, 1), device_ms=round(exact32_dev_ms, 1),
                   note="the unfused complex-f32 path (one exact kernel per gate, QuantumStateF32 arithmetic) vs "
                        "the exact complex-f64 run")
        err["exact_c32"] = e32
        del a32, a64, a32x
        log("f32 error: %s" % err)
    out["f32_error_vs_f64"] = err

    dense = None
    if prec == "c32" and not args.no_dense:
        log("dense 5-qubit blocks on the matrix cores ...")
        dms = []
        for k in range(2):
            rc = L.rq_run_circuit(st.handle, ptr(arr), ngates, None, 0, None, 0, C.byref(nout), 2, C.byref(met))
            if rc:
                raise RuntimeError(L.rq_last_error().decode())
            dms.append(met.device_time_ms)
        blocks = int(met.passes)
        per_block_ms = min(dms) / blocks
        flops = 256.0 * n  # real 64x64 GEMM row per 32 amplitudes: 8 * 32 flop per amplitude
        dense = {"kernel": "k_dense5_f32", "blocks": blocks, "circuit_ms": round(min(dms), 3),
                 "gate_apps_per_s": round(ngates / (min(dms) * 1e-3), 2),
                 "mean_block_ms": round(per_block_ms, 4),
                 "mfma_TFLOPs": round(flops / (per_block_ms * 1e-3) / 1e12, 2),
                 "mfma_peak_TFLOPs": MFMA_F32_PEAK_TFLOPS,
                 "mfma_frac": round(flops / (per_block_ms * 1e-3) / 1e12 / MFMA_F32_PEAK_TFLOPS, 4),
                 "hbm_GBps": round(stage_bytes / (per_block_ms * 1e-3) / 1e9, 1),
                 "note": "same circuit as dense 32x32 complex unitaries (v_mfma_f32_32x32x2_f32); "
                         "the fused tile schedule needs %d passes instead of %d" % (passes, blocks)}
        log("dense: %s" % dense)
    out["dense_blocks"] = dense
    st.close()
    return out


def c64_32q(args, L):
    """The same generator at 32 qubits complex f64, depth 20 (a 64 GiB state)."""
    from ruqu_amd import QuantumState, workloads
    from ruqu_amd._native import Metrics, ptr
    nq = 32
    arr = workloads.random_circuit(nq, 20, args.seed).to_array()
    log("32q c64 depth 20 (%d gates) ..." % len(arr))
    st = QuantumState.new_with_seed(nq, args.seed, "c64")
    met, nout = Metrics(), C.c_size_t(0)
    ms = []
    for k in range(2):
        if L.rq_run_circuit(st.handle, ptr(arr), len(arr), None, 0, None, 0, C.byref(nout), 1, C.byref(met)):
            raise RuntimeError(L.rq_last_error().decode())
        ms.append(met.device_time_ms)
    st.close()
    stage_ms = ms[-1] / met.passes
    r = {"qubits": nq, "depth": 20, "gates": len(arr), "stages": int(met.passes), "circuit_ms": round(ms[-1], 2),
         "gate_apps_per_s": round(len(arr) / (ms[-1] * 1e-3), 1), "mean_stage_ms": round(stage_ms, 3),
         "stage_hbm_GBps": round(2.0 * (1 << nq) * 16 / (stage_ms * 1e-3) / 1e9, 1)}
    log("32q c64: %s" % r)
    return r


# ---- configuration 1: 20 qubits complex f64 ------------------------------------

def config1(args, L):
    from ruqu_amd import QuantumState, SimConfig, Simulator, workloads
    n, depth, shots = 20, 20, 10_000
    c = workloads.random_circuit(n, depth, 1)
    arr = c.to_array()
    draws = workloads.uniforms(workloads.shot_seed(1), shots * n).reshape(shots, n)
    log("config 1: %d qubits c64 depth %d (%d gates), %d shots" % (n, depth, len(arr), shots))

    def gpu_once():
        t0 = time.perf_counter()
        r = Simulator.run_with_config(c, SimConfig(precision="c64"))
        t1 = time.perf_counter()
        amps = r.state.state_vector()
        t2 = time.perf_counter()
        z = r.state.expectation_z_all()
        t3 = time.perf_counter()
        bits, ties = r.state.sample_shots(draws)
        t4 = time.perf_counter()
        r.state.close()
        return (t1 - t0, t2 - t1, t3 - t2, t4 - t3, t4 - t0), amps, z, bits

    # steady state: one untimed run, then the run with the median total of
    # three (the first host reads of a process pay the runtime's set-up and
    # first-touch page faults, up to 10 ms for the 16 MB state)
    gpu_once()
    runs = sorted((gpu_once() for _ in range(3)), key=lambda x: x[0][4])
    times, amps, z, bits = runs[1]
    fz = Simulator.run_with_config(c, SimConfig(precision="c64", fusion=True))
    t0 = time.perf_counter()
    fz2 = Simulator.run_with_config(c, SimConfig(precision="c64", fusion=True))
    fused_s = time.perf_counter() - t0
    fz.state.close()
    fz2.state.close()
    out = {"workload": "random circuit 20 qubits complex-f64 depth 20, |0..0>, all amplitudes, <Z_i> for every "
                       "qubit, 10^4 shots (measure_all semantics, fed draws)",
           "gates": len(arr),
           "gpu_exact_s": {"circuit": round(times[0], 5), "read_amplitudes": round(times[1], 5),
                           "z_all": round(times[2], 5), "shots_10k": round(times[3], 5), "total": round(times[4], 5)},
           "gpu_exact_gate_apps_per_s": round(len(arr) / times[0], 1),
           "gpu_fused_circuit_s": round(fused_s, 5)}
    if not args.no_cpu:
        th = cpu_threads()
        done, s = cpu_gates(n, arr, args.cpu_budget / 2, "scalar")
        circ_scalar = s / done * len(arr)
        done2, s2 = cpu_gates(n, arr, args.cpu_budget / 2, "simd", th)
        circ_simd = s2 / done2 * len(arr)
        per_shot = cpu_measure_all(n, 3.0, amps)
        out["cpu"] = {
            "scalar_circuit_s": round(circ_scalar, 3), "simd_circuit_s": round(circ_simd, 3), "simd_threads": th,
            "measure_all_per_shot_s": round(per_shot, 5),
            "run_shots_as_written_s": round(shots * (circ_simd + per_shot), 1),
            "run_shots_measure_all_on_copies_s": round(circ_simd + shots * per_shot, 2),
            "sample": "scalar: first %d gates in %.1f s; simd.rs path: first %d gates in %.1f s (%d threads); "
                      "measure_all timed on copies; totals extrapolated to %d gates and %d shots"
                      % (done, s, done2, s2, th, len(arr), shots),
            "kind": "port"}
        gpu_total = times[4]
        out["speedup_vs_cpu_simd"] = {
            "circuit": round(circ_simd / times[0], 1),
            "with_shots_measure_all_on_copies": round((circ_simd + shots * per_shot) / gpu_total, 1)}
    log("config 1: %s" % {k: v for k, v in out.items() if k != "workload"})
    return out


# ---- configuration 3: QAOA MaxCut n = 26, p = 4 ----------------------------------

def config3(args, L):
    from ruqu_amd import workloads
    from ruqu_amd.qaoa import Graph, QaoaInstance, build_qaoa_circuit
    n, p, shots = 26, 4, 100_000
    edges = workloads.random_regular_graph(n, 3, 3)
    u = workloads.uniforms(workloads.stream_seed(3, 1), 2 * p)
    gammas, betas = list(u[:p] * math.pi), list(u[p:] * math.pi)
    g = Graph.unweighted(n, edges)
    draws = workloads.uniforms(workloads.shot_seed(3), shots * n).reshape(shots, n)
    log("config 3: QAOA n=%d p=%d (3-regular, seed 3), %d shots" % (n, p, shots))
    inst = QaoaInstance(g, "c64")
    inst.cost(gammas, betas, "exact")
    t0 = time.perf_counter()
    e_exact = inst.cost(gammas, betas, "exact")
    t1 = time.perf_counter()
    e_fused = inst.cost(gammas, betas, "fused")
    t2 = time.perf_counter()
    inst.sample_cuts(draws)  # untimed: the first sampling of a process pays the runtime's set-up (about 20 ms)
    t2s = time.perf_counter()
    bits, cuts, ties = inst.sample_cuts(draws)
    t3 = time.perf_counter()
    del inst
    out = {"workload": "QAOA MaxCut, random 3-regular graph n=26 (seed 3), unit weights, p=4, angles U(0,pi), "
                       "complex f64: <H_C> and 10^5 sampled cuts",
           "expected_cut": e_exact, "expected_cut_fused": e_fused, "sampled_mean_cut": float(cuts.mean()),
           "best_sampled_cut": int(cuts.max()), "near_ties": int(ties),
           "gpu_s": {"cost_exact": round(t1 - t0, 5), "cost_fused": round(t2 - t1, 5),
                     "shots_100k": round(t3 - t2s, 5), "total_exact_path": round((t1 - t0) + (t3 - t2s), 5)}}
    if not args.no_cpu:
        from oracle import oracle as orc
        arr = build_qaoa_circuit(g, gammas, betas).to_array()
        th = cpu_threads()
        done, s = cpu_gates(n, arr, args.cpu_budget / 2, "scalar")
        done2, s2 = cpu_gates(n, arr, args.cpu_budget / 2, "simd", th)
        # one <Z_i Z_j> pass of the reference's cost (qaoa.rs:192-198), timed on a state of the same size
        a = np.zeros(1 << n, dtype=np.complex128)
        a[0] = 1
        t4 = time.perf_counter()
        orc.expectation(a, n, [(edges[0][0], 3), (edges[0][1], 3)])
        zz_s = time.perf_counter() - t4
        del a
        cost_scalar = s / done * len(arr) + zz_s * len(edges)
        cost_simd = s2 / done2 * len(arr) + zz_s * len(edges)
        out["cpu"] = {"scalar_cost_s": round(cost_scalar, 2), "simd_cost_s": round(cost_simd, 2), "simd_threads": th,
                      "sample": "scalar: first %d of %d gates in %.1f s; simd.rs path: first %d gates in %.1f s "
                                "(%d threads); one <ZZ> pass %.2f s; extrapolated to the full cost evaluation"
                                % (done, len(arr), s, done2, s2, th, zz_s), "kind": "port"}
        out["speedup_vs_cpu_simd_cost"] = round(cost_simd / (t1 - t0), 1)
    log("config 3: %s" % {k: v for k, v in out.items() if k != "workload"})
    return out


# ---- configuration 4: batched QAOA + VQE -----------------------------------------

def config4(args, L):
    from ruqu_amd import config4 as cfg4
    from ruqu_amd.batch import Batch
    log("config 4: building 4096 QAOA + 512 VQE instances ...")
    qa = cfg4.qaoa_instances()
    vq = cfg4.vqe_instances()
    b = Batch()
    for n, e, w, p in qa:
        b.add_qaoa(n, e, w, p)
    for n, layers, h in vq:
        b.add_vqe(n, layers, h)
    total = len(qa) + len(vq)
    b.evaluate()
    t0 = time.perf_counter()
    reps = 3
    for _ in range(reps):
        b.evaluate()
    ev_s = (time.perf_counter() - t0) / reps
    b.optimize(1)  # captures the iteration's graph
    iters = 3
    t1 = time.perf_counter()
    dev_ms = b.optimize(iters)
    it_s = (time.perf_counter() - t1) / iters
    out = {"workload": "4096 QAOA MaxCut (n~U{10..22}, G(n,0.5), U(0,1) weights, p~U{1..5}) + 512 VQE (n~U{12..16}, "
                       "Ry-CZ ansatz 4 layers, 100 Pauli strings of weight <= 4, N(0,1) coefficients), complex-f32 "
                       "states, FP64 reductions, one persistent batch, one hipGraph per optimizer iteration",
           "instances": total, "evaluate_s": round(ev_s, 5), "evaluate_inst_per_s": round(total / ev_s, 1),
           "iteration_s": round(it_s, 4), "iteration_device_ms": round(dev_ms / iters, 2),
           "iteration_inst_per_s": round(total / it_s, 1), "tile": b.tile()}
    del b
    if not args.no_cpu:
        # the reference's evaluate_qaoa_cost (qaoa.rs:179-201) on the scalar path, a sample of the same instances
        from oracle import oracle as orc
        t2 = time.perf_counter()
        k = 0
        amps_done = 0
        while k < len(qa) and time.perf_counter() - t2 < args.cpu_budget / 2:
            n, e, w, p = qa[k]
            orc.qaoa_cost(n, e, w, [0.5] * p, [0.5] * p)
            amps_done += (1 << n)
            k += 1
        s = time.perf_counter() - t2
        out["cpu"] = {"evaluate_inst_per_s": round(k / s, 2), "kind": "port",
                      "sample": "the first %d QAOA instances on the restated scalar path (one thread), %.1f s; the "
                                "sample's mean state is %.0f amplitudes against %.0f for the whole QAOA set"
                                % (k, s, amps_done / k, float(np.mean([1 << q[0] for q in qa])))}
    log("config 4: %s" % {k: v for k, v in out.items() if k != "workload"})
    return out


# ---- configuration 5: learned warm start -------------------------------------------

def config5(args):
    params_path = os.path.join(ROOT, "ruqu_amd", "data", "warmstart_params.npy")
    if not os.path.exists(params_path):
        log("qaoa part skipped: no trained parameters at %s" % params_path)
        return None
    from ruqu_amd import warmstart
    from ruqu_amd.mpnn import Network
    log("qaoa: %d held-out instances, plain %d iterations vs warm start" % (args.qaoa, args.qaoa_iters))
    insts = warmstart.held_out(args.qaoa)
    warmstart.run_plain(insts[:32], 1)  # warm-up of the batched kernels
    net = Network.load(params_path)
    base = warmstart.plain_baseline(insts, args.qaoa_iters)
    cands = [net, warmstart.ConstantPredictor.load()]
    r = warmstart.evaluate(cands, insts, args.qaoa_iters, 0, log=log, plain=base, rescue_iters=args.qaoa_rescue)
    out = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
    out["workload"] = ("held-out MaxCut: G(n,0.5) n~U{10..22} U(0,1) weights p~U{1..8}, complex-f32 states, "
                       "FP64 reductions")
    return out


def wide_pass_case(spec):
    """--wide-pass-case n:p:count[:tile[:exact]]: `count` QAOA MaxCut
    instances of n qubits and depth p (G(n, 0.5), U(0,1) weights, angles
    0.3 / 0.7) evaluated three times on the batched engine, so that every
    k_wide_pass launch covers states of one size: per-pass traffic counters
    (pass s of an evaluation is launch s mod (p + 1)).  Prints one JSON line."""
    from ruqu_amd import workloads
    from ruqu_amd.batch import Batch
    f = spec.split(":")
    n, p, count = int(f[0]), int(f[1]), int(f[2])
    tile = int(f[3]) if len(f) > 3 else None
    exact = len(f) > 4 and f[4] == "exact"
    b = Batch()
    for i in range(count):
        e, w = workloads.erdos_renyi_graph(n, 0.5, 1000 + i)
        b.add_qaoa(n, e, w, p, [0.3] * p, [0.7] * p)
    if tile:
        b.set_tile(tile, exact=exact)
    ms = []
    for _ in range(3):
        en, _ = b.evaluate()
        ms.append(b.last_ms())
    print(json.dumps({"wide_pass_case": spec, "qubits": n, "p": p, "instances": count, "tile": b.tile(),
                      "exact": exact, "evaluate_ms": [round(x, 3) for x in ms], "energy0": float(en[0])}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--qubits", type=int, default=30)
    ap.add_argument("--depth", type=int, default=40)
    ap.add_argument("--precision", default="c32")
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=16.0, help="seconds of CPU work per CPU leg")
    ap.add_argument("--qaoa", type=int, default=4096, help="held-out QAOA instances of the warm start (0: skip)")
    ap.add_argument("--no-f64-check", action="store_true", help="skip the f32-vs-exact-f64 error of the circuit")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense 5-qubit block (matrix-core) run")
    ap.add_argument("--only-headline", action="store_true", help="the headline circuit alone (no other config)")
    ap.add_argument("--qaoa-iters", type=int, default=16)
    ap.add_argument("--qaoa-rescue", type=int, default=6,
                    help="plain updates from the warm angles for instances below the 1%% bar before falling back")
    ap.add_argument("--wide-pass-case", default=None, metavar="N:P:COUNT[:TILE[:exact]]",
                    help="only evaluate a uniform batch of QAOA instances (traffic profiles of the batched passes)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU share "
              "(multi-GPU qubit sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    if args.wide_pass_case:
        return wide_pass_case(args.wide_pass_case)
    from ruqu_amd._native import lib
    L = lib()
    h = headline(args, L)
    ngates, el = h["ngates"], h["el"]

    cpu = cpu_f32 = None
    if not args.no_cpu:
        log("cpu baseline (simd.rs path, complex f64) ...")
        th = cpu_threads()
        done, s = cpu_gates(args.qubits, h["arr"], args.cpu_budget * 0.75, "simd", th)
        cpu = {"value": done / s, "unit": "gate-apps/s", "cores": th, "kind": "port",
               "sample": "first %d gates of the %d-qubit circuit on the restated simd.rs *_best path (the reference's "
                         "fastest CPU path: AVX2 + thread pool, complex f64), %d threads, %.1f s"
                         % (done, args.qubits, th, s)}
        log("cpu f32 scalar baseline ...")
        done, s = cpu_gates(args.qubits, h["arr"], args.cpu_budget * 0.75, "scalar", 1, f32=True)
        cpu_f32 = {"value": done / s, "unit": "gate-apps/s", "cores": 1, "kind": "port",
                   "sample": "first %d gates of the %d-qubit c32 circuit on the restated QuantumStateF32 scalar path "
                             "(mixed_precision.rs:396-458), %.1f s" % (done, args.qubits, s)}

    extra = {}
    if not args.only_headline and args.qubits == 30:
        extra["c64_32q"] = c64_32q(args, L)
        extra["config1"] = config1(args, L)
        extra["config3"] = config3(args, L)
        extra["config4"] = config4(args, L)
    qaoa = config5(args) if (args.qaoa > 0 and not args.only_headline) else None

    prec, nq = args.precision, args.qubits
    line = {
        "metric": "gate-apps/s (%dq %s random circuit, depth %d)" % (nq, prec, args.depth),
        "value": round(ngates * args.steps / el, 2),
        "unit": "gate-apps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32" if prec == "c32" else "f64",
        "data": "synthetic (SplitMix64 seed %d random circuit)" % args.seed,
        "config": {"workload": "random circuit %d qubits complex-%s depth %d, %d gates, fused schedule" %
                               (nq, "f32" if prec == "c32" else "f64", args.depth, ngates),
                   "qubits": nq, "depth": args.depth, "gates": ngates, "passes_per_step": h["passes"]},
        "roofline": h["roofline"],
        "roofline_valu": h.get("roofline_valu"),
        "cpu_baseline": cpu,
        "cpu_baseline_f32_scalar": cpu_f32,
        "per_gate_1q_GBps": h["per_gate_1q_GBps"],
        "per_gate_1q_min_GBps": h["per_gate_1q_min_GBps"],
        "copy_GBps": h["copy_GBps"],
        "f32_error_vs_f64": h["f32_error_vs_f64"],
        "dense_blocks": h["dense_blocks"],
        "qaoa": qaoa,
    }
    line.update(extra)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
