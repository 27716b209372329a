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
                of the same circuit.
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

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(nq, arr, args.cpu_budget)

    value = ngates * args.steps / el
    line = {
        "metric": "gate-apps/s (30q c32 random circuit, depth 40)",
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
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel": "k_fused_stage", "bytes_per_launch": stage_bytes,
                     "mean_launch_ms": round(mean_stage_ms, 4)},
        "per_gate_1q_GBps": per_gate,
        "copy_GBps": copy_gbs,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
