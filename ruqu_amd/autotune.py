"""Autotuner for the launch shapes of the device path.

Each tuner times the real kernels on the current device (HIP events), keeps
the fastest candidate, applies it, and records it in a JSON cache keyed by
the device name and the workload's signature, so the next process applies
the cached choice without timing again:

  * tune_streaming: amplitude groups per thread (1, 2, 4, 8) and
    non-temporal access of the exact per-gate and copy kernels
    (rq_tune_set keys 0 and 1).
  * tune_fused: tile qubits m and low qubits c of the fused schedule
    (keys 2 and 3), then for complex-f32 states the matrix-core threshold
    (key 4), on a random circuit of the caller's size.
  * tune_batch: LDS tile width of the batched engine's states above 13
    qubits (rq_batch_set_tile), on the caller's own batch.

Cache file: $RUQU_AMD_TUNE_CACHE, else ~/.cache/ruqu_amd/tune.json.  The
reference has no autotuner; the knobs only change launch shapes, never
results beyond floating-point rounding order.
"""
import ctypes as C
import hashlib
import json
import os

import numpy as np

from . import workloads
from ._native import lib
from .state import QuantumState, SimConfig, Simulator, check

KEY_UNROLL, KEY_NONTEMPORAL, KEY_FUSED_M, KEY_FUSED_C, KEY_FUSED_MF = 0, 1, 2, 3, 4


def cache_path():
    return os.environ.get("RUQU_AMD_TUNE_CACHE", os.path.join(os.path.expanduser("~"), ".cache", "ruqu_amd",
                                                               "tune.json"))


def _load():
    try:
        with open(cache_path()) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def _store(key, value):
    d = _load()
    d[key] = value
    p = cache_path()
    try:
        os.makedirs(os.path.dirname(p), exist_ok=True)
        tmp = p + ".tmp%d" % os.getpid()
        with open(tmp, "w") as f:
            json.dump(d, f, indent=1, sort_keys=True)
        os.replace(tmp, p)
    except OSError:
        pass  # an unwritable cache only costs a re-tune


def _device():
    import torch
    return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"


def _key(kind, sig):
    return "%s|%s|%s" % (_device(), kind, sig)


def set_knob(key, value):
    check(lib().rq_tune_set(key, int(value)))


def get_knob(key):
    return int(lib().rq_tune_get(key))


def tune_streaming(num_qubits=28, precision="c32", reps=10, use_cache=True):
    """Picks (groups per thread, non-temporal) for the exact 1q kernel,
    timed on qubits 1, num_qubits/2 and num_qubits-1 of a fresh state."""
    key = _key("streaming", "%s-%d" % (precision, num_qubits))
    cached = _load().get(key) if use_cache else None
    if cached:
        set_knob(KEY_UNROLL, cached["unroll"])
        set_knob(KEY_NONTEMPORAL, cached["nontemporal"])
        return cached
    st = QuantumState.new(num_qubits, precision)
    m = np.ascontiguousarray(np.array([[1, 1], [1, -1]]) / np.sqrt(2), dtype=np.complex128)  # H
    times = {}
    for u in (1, 2, 4, 8):
        for nt in (0, 1):
            set_knob(KEY_UNROLL, u)
            set_knob(KEY_NONTEMPORAL, nt)
            tot = 0.0
            for q in (1, num_qubits // 2, num_qubits - 1):
                ms = C.c_double()
                check(lib().rq_time_1q(st.handle, q, m.ctypes.data_as(C.c_void_p), reps, C.byref(ms)))
                tot += ms.value
            times[(u, nt)] = tot
    (u, nt), best = min(times.items(), key=lambda kv: kv[1])
    set_knob(KEY_UNROLL, u)
    set_knob(KEY_NONTEMPORAL, nt)
    res = {"unroll": u, "nontemporal": nt, "ms_sum": best,
           "tried": {"%d,%d" % k: v for k, v in times.items()}}
    _store(key, res)
    return res


def tune_fused(num_qubits, precision="c32", depth=4, seed=7, candidates=None, mf_candidates=(-1, 512, 768, 1024),
               use_cache=True):
    """Picks the fused schedule's (m, c), and for complex-f32 states its
    matrix-core threshold, by timing a random circuit of the caller's width
    (depth layers) through rq_run_circuit(fusion=1)."""
    key = _key("fused", "%s-%d" % (precision, num_qubits))
    cached = _load().get(key) if use_cache else None
    if cached:
        set_knob(KEY_FUSED_M, cached["m"])
        set_knob(KEY_FUSED_C, cached["c"])
        if "mf" in cached:
            set_knob(KEY_FUSED_MF, cached["mf"])
        return cached
    if candidates is None:
        candidates = [(m, c) for m in (10, 11, 12) for c in (3, 4, 5, 6) if c < m and m <= num_qubits]
    circ = workloads.random_circuit(num_qubits, depth, seed)
    st = QuantumState.new(num_qubits, precision)
    cfg = SimConfig(precision=precision, fusion=True)
    times = {}
    for m, c in candidates:
        check(lib().rq_fusion_configure(m, c))
        Simulator.run_with_config(circ, cfg, state=st)  # warm-up / plan
        r = Simulator.run_with_config(circ, cfg, state=st)
        times[(m, c)] = r.metrics.device_time_ms
    (m, c), best = min(times.items(), key=lambda kv: kv[1])
    check(lib().rq_fusion_configure(m, c))
    res = {"m": m, "c": c, "ms": best, "tried": {"%d,%d" % k: v for k, v in times.items()}}
    if precision == "c32" and mf_candidates:
        mf_times = {}
        for mf in mf_candidates:
            set_knob(KEY_FUSED_MF, mf)
            Simulator.run_with_config(circ, cfg, state=st)
            mf_times[mf] = Simulator.run_with_config(circ, cfg, state=st).metrics.device_time_ms
        mf = min(mf_times, key=mf_times.get)
        set_knob(KEY_FUSED_MF, mf)
        res.update(mf=mf, ms=mf_times[mf], tried_mf={str(k): v for k, v in mf_times.items()})
    _store(key, res)
    return res


def batch_signature(batch):
    """Histogram of (kind, qubits, layers) over the batch's instances."""
    h = {}
    for kind, n, p in zip(batch._kinds, batch._n, batch._layers):
        h[(kind, n, p)] = h.get((kind, n, p), 0) + 1
    txt = ";".join("%s,%d,%d:%d" % (k[0], k[1], k[2], v) for k, v in sorted(h.items()))
    return hashlib.sha1(txt.encode()).hexdigest()[:16]


def tune_batch(batch, candidates=(10, 11, 12, 13), repeats=2, use_cache=True):
    """Picks the batched engine's LDS tile width for states above 13 qubits
    by timing rq_batch_evaluate on the batch itself."""
    key = _key("batch", batch_signature(batch))
    cached = _load().get(key) if use_cache else None
    if cached:
        batch.set_tile(cached["tile"])
        return cached
    nmax = max([n for n in batch._n if n > 13], default=0)
    times = {}
    for tq in candidates:
        if 2 * tq < nmax:
            continue
        batch.set_tile(tq)
        batch.evaluate()  # plan upload, warm-up
        best = float("inf")
        for _ in range(repeats):
            batch.evaluate()
            best = min(best, batch.last_ms())
        times[tq] = best
    tq = min(times, key=times.get)
    batch.set_tile(tq)
    res = {"tile": tq, "ms": times[tq], "tried": {str(k): v for k, v in times.items()}}
    _store(key, res)
    return res
