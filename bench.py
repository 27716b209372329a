"""Throughput of the state-vector engine on the 30-qubit complex-f32 random
circuit (depth 40; Rx, Ry, Rz on every qubit then 10 CNOT/CZ pairs per layer),
on one MI355X.

A step = one full simulation of the circuit from |0..0> through the fused
schedule (every gate applied; the state stays in HBM).  `value` counts the
circuit's original gate applications per second.  The line also carries:
  roofline      the fused stage kernel: algorithmic bytes (2 * 2^30 * 8 B per
                stage) / mean stage duration from HIP events in the timed
                region, against the 8 TB/s HBM peak;
  per_gate      the exact single-qubit gate kernel on the same state (one
                HBM pass per gate), GB/s per target qubit;
  cpu_baseline  the reference's f32 CPU path (QuantumStateF32 scalar loops,
                restated in the oracle), one host thread, on the first gates
                of the same circuit;
  cpu_baseline_simd  the reference's fastest CPU path beside it: simd.rs's
                AVX2 + thread-pool kernels on a complex-f64 state (restated
                in the oracle), up to 16 host threads, same leading gates;
  qaoa          the batched QAOA part of the metric: held-out MaxCut
                instances (G(n, 0.5), n ~ U{10..22}, U(0,1) weights,
                p ~ U{1..8}) through the plain batched optimizer (run_qaoa's
                rule, 16 iterations) and through the learned warm start
                (network prediction and per-depth constant, each judged by
                one batched evaluation; instances whose
                expected cut is below 99 % of plain's get 6 plain updates
                from their best angles, and those still below fall back to
                the plain optimizer): instances/s of both and the fraction
                within 1 %.
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(nq, arr, budget_s=20.0):
    """Oracle f32 scalar path (mixed_precision.rs:396-458), 1 thread, on the
    leading gates of the circuit until `budget_s` of work is done."""
    from oracle import oracle as orc
    import ctypes as C
    amps = np.zeros(1 << nq, dtype=np.complex64)
    amps[0] = 1
    lib = orc.lib()
    p = amps.ctypes.data_as(C.c_void_p)
    done = 0
    t0 = time.perf_counter()
    cur = C.c_size_t(0)
    nout = C.c_size_t(0)
    while done < len(arr) and time.perf_counter() - t0 < budget_s:
        g = np.ascontiguousarray(arr[done:done + 1]).view(orc.GATE_DTYPE)
        rc = lib.orc_apply_gate(p, 1, nq, g.ctypes.data_as(C.c_void_p), None, 0, C.byref(cur), None,
                                C.byref(nout))
        if rc:
            raise RuntimeError("oracle gate failed")
        done += 1
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "gate-apps/s", "cores": 1, "kind": "port",
            "sample": "first %d gates of the %d-qubit c32 circuit on the restated QuantumStateF32 scalar path "
                      "(mixed_precision.rs:396-458), %.1f s" % (done, nq, el)}


def cpu_simd_baseline(nq, arr, budget_s=12.0):
    """Oracle restatement of simd.rs's *_best kernels (simd.rs:270-377: AVX2
    with a thread pool over a complex-f64 state, the reference's fastest CPU
    path), on the leading gates of the same circuit, with the host threads
    this process may use (at most 16)."""
    from oracle import oracle as orc
    threads = max(1, min(len(os.sched_getaffinity(0)), 16))
    amps = np.zeros(1 << nq, dtype=np.complex128)
    amps[0] = 1
    two = (orc.CNOT, orc.CZ, orc.SWAP, orc.RZZ)
    done = 0
    t0 = time.perf_counter()
    while done < len(arr) and time.perf_counter() - t0 < budget_s:
        g = np.ascontiguousarray(arr[done:done + 1]).view(orc.GATE_DTYPE)[0]
        if int(g["kind"]) in two:
            orc.simd_apply_2q_best(amps, nq, int(g["q0"]), int(g["q1"]), orc.matrix_2q(g), threads)
        else:
            orc.simd_apply_1q_best(amps, nq, int(g["q0"]), orc.matrix_1q(g), threads)
        done += 1
    el = time.perf_counter() - t0
    return {"value": done / el, "unit": "gate-apps/s", "cores": threads, "kind": "port",
            "sample": "first %d gates of the %d-qubit circuit on the restated simd.rs *_best path (complex f64, "
                      "AVX2, %d threads), %.1f s" % (done, nq, threads, el)}


def measured_traffic(path):
    """HBM bytes per launch of the dominant kernel, from the calibrated
    FETCH_SIZE / WRITE_SIZE profile summary (None when absent)."""
    try:
        for line in open(path):
            if line.startswith("traffic_bytes_per_launch"):
                return float(line.split("=")[1])
    except OSError:
        pass
    return None


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
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--qaoa", type=int, default=1024, help="held-out QAOA instances (0: skip)")
    ap.add_argument("--no-f64-check", action="store_true", help="skip the f32-vs-f64 error of the circuit")
    ap.add_argument("--no-dense", action="store_true", help="skip the dense 5-qubit block (matrix-core) run")
    ap.add_argument("--qaoa-iters", type=int, default=16)
    ap.add_argument("--qaoa-rescue", type=int, default=6,
                    help="plain updates from the warm angles for instances below the 1%% bar before falling back")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py runs on one GPU only: this build measures the single-GPU share "
              "(multi-GPU qubit sharding is a design note, not implemented)", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    from ruqu_amd import QuantumState
    from ruqu_amd._native import Metrics, lib, ptr
    from ruqu_amd import workloads

    nq, prec = args.qubits, args.precision
    amp_bytes = 8 if prec == "c32" else 16
    log("building circuit %dq depth %d" % (nq, args.depth))
    circ = workloads.random_circuit(nq, args.depth, args.seed)
    arr = circ.to_array()
    ngates = len(arr)
    st = QuantumState.new_with_seed(nq, args.seed, prec)
    L = lib()
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

    # exact single-qubit kernel bandwidth on the same state (per target qubit)
    per_gate = {}
    m = np.array([[0.6, 0.8j], [0.8j, 0.6]], dtype=np.complex128)
    for q in sorted(set([0, 1, 4, 5, 12, nq // 2, nq - 2, nq - 1])):
        avg = C.c_double()
        L.rq_time_1q(st.handle, q, ptr(m), 1, C.byref(avg))  # warm
        rc = L.rq_time_1q(st.handle, q, ptr(m), 5, C.byref(avg))
        if rc:
            raise RuntimeError(L.rq_last_error().decode())
        per_gate[str(q)] = round(stage_bytes / (avg.value * 1e-3) / 1e9, 1)
    cavg = C.c_double()
    copy_gbs = None
    if L.rq_time_copy(st.handle, 5, C.byref(cavg)) == 0:
        copy_gbs = round(stage_bytes / (cavg.value * 1e-3) / 1e9, 1)
    log("per-gate GB/s %s, copy %s GB/s" % (per_gate, copy_gbs))

    err = None
    if prec == "c32" and not args.no_f64_check and nq <= 31:
        log("f64 run of the same circuit for the f32 error ...")
        step()  # the per-gate timing above changed the state: run the circuit again
        a32 = st.state_vector()
        st64 = QuantumState.new_with_seed(nq, args.seed, "c64")
        rc = L.rq_run_circuit(st64.handle, ptr(arr), ngates, None, 0, None, 0, C.byref(nout), 1, C.byref(met))
        if rc:
            raise RuntimeError(L.rq_last_error().decode())
        a64 = st64.state_vector()
        del st64
        mx, d2, n64, n32 = 0.0, 0.0, 0.0, 0.0
        ip = 0j
        ch = 1 << 24
        for o in range(0, n, ch):
            x = a32[o:o + ch].astype(np.complex128)
            y = a64[o:o + ch]
            d = x - y
            mx = max(mx, float(np.max(np.abs(d))))
            d2 += float(np.vdot(d, d).real)
            n64 += float(np.vdot(y, y).real)
            n32 += float(np.vdot(x, x).real)
            ip += complex(np.vdot(y, x))
        del a32, a64
        err = {"max_abs": mx, "rel_l2": (d2 / n64) ** 0.5, "one_minus_fidelity": 1.0 - abs(ip) ** 2 / (n32 * n64),
               "note": "complex-f32 fused run vs complex-f64 fused run of the same circuit on the device "
                       "(the f64 path matches the FP64 CPU oracle to 1e-12)"}
        log("f32 error: %s" % err)

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
                 "mfma_peak_TFLOPs": 157.3,
                 "mfma_frac": round(flops / (per_block_ms * 1e-3) / 1e12 / 157.3, 4),
                 "hbm_GBps": round(stage_bytes / (per_block_ms * 1e-3) / 1e9, 1),
                 "note": "same circuit as dense 32x32 complex unitaries (v_mfma_f32_32x32x2_f32); "
                         "the fused tile schedule above needs %d passes instead of %d" % (passes, blocks)}
        log("dense: %s" % dense)

    cpu = cpu_simd = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(nq, arr, args.cpu_budget)
        log("cpu simd baseline ...")
        cpu_simd = cpu_simd_baseline(nq, arr, min(args.cpu_budget, 12.0))

    qaoa = None
    params_path = os.path.join(ROOT, "ruqu_amd", "data", "warmstart_params.npy")
    if args.qaoa > 0 and os.path.exists(params_path):
        from ruqu_amd import warmstart
        from ruqu_amd.mpnn import Network
        del st  # free the 8 GiB state before the batched instances
        log("qaoa: %d held-out instances, plain %d iterations vs warm start" % (args.qaoa, args.qaoa_iters))
        insts = warmstart.held_out(args.qaoa)
        warmstart.run_plain(insts[:32], 1)  # warm-up of the batched kernels
        net = Network.load(params_path)
        base = warmstart.plain_baseline(insts, args.qaoa_iters)
        cands = [net, warmstart.ConstantPredictor.load()]
        r = warmstart.evaluate(cands, insts, args.qaoa_iters, 0, log=log, plain=base, rescue_iters=args.qaoa_rescue)
        qaoa = {k: (round(v, 4) if isinstance(v, float) else v) for k, v in r.items()}
        qaoa["workload"] = ("held-out MaxCut: G(n,0.5) n~U{10..22} U(0,1) weights p~U{1..8}, complex-f32 states, "
                            "FP64 reductions")
    elif args.qaoa > 0:
        log("qaoa part skipped: no trained parameters at %s" % params_path)

    value = ngates * args.steps / el
    line = {
        "metric": "gate-apps/s (%dq %s random circuit, depth %d)" % (nq, prec, args.depth),
        "value": round(value, 2),
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
                   "qubits": nq, "depth": args.depth, "gates": ngates, "passes_per_step": passes},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": (measured_traffic(os.path.join(ROOT, "profiles", "fused_stage_30q_c32_hbm_traffic.md"))
                                 if (nq, prec, args.depth) == (30, "c32", 40) else None),
                     "kernel": "k_fused_stage", "bytes_per_launch": stage_bytes,
                     "mean_launch_ms": round(mean_stage_ms, 4)},
        "per_gate_1q_GBps": per_gate,
        "copy_GBps": copy_gbs,
        "cpu_baseline": cpu,
        "cpu_baseline_simd": cpu_simd,
        "f32_error_vs_f64": err,
        "dense_blocks": dense,
        "qaoa": qaoa,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
