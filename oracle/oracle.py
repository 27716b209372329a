"""ORACLE — test infrastructure only.

ctypes wrapper of oracle/liboracle.so, the CPU restatement of ruqu-core's
scalar path (state.rs, mixed_precision.rs, simulator.rs, gate.rs,
optimizer.rs), its simd.rs kernels, and the ruqu-algorithms QAOA/VQE
drivers.  Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of
bench.py may import this module; it is the checker, never the product.
"""
import contextlib
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

GATE_DTYPE = np.dtype(
    [("kind", "<i4"), ("q0", "<u4"), ("q1", "<u4"), ("_pad", "<u4"), ("theta", "<f8"), ("u", "<f8", (8,))]
)
OUTCOME_DTYPE = np.dtype([("qubit", "<u4"), ("result", "<i4"), ("probability", "<f8")])

# gate.rs:8-36 order
H, X, Y, Z, S, SDG, T, TDG, RX, RY, RZ, PHASE, CNOT, CZ, SWAP, RZZ, MEASURE, RESET, BARRIER, UNITARY1Q = range(20)


class QaoaResult(C.Structure):
    _fields_ = [("best_cut_value", C.c_double), ("iterations", C.c_uint32), ("converged", C.c_int32),
                ("best_index", C.c_uint64)]


class VqeResult(C.Structure):
    _fields_ = [("optimal_energy", C.c_double), ("num_iterations", C.c_uint32), ("converged", C.c_int32)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def load(path):
    """ctypes handle of an oracle build at `path`, with every signature set."""
    handle = C.CDLL(path)
    vp, u32, sz, d, i = C.c_void_p, C.c_uint32, C.c_size_t, C.c_double, C.c_int
    sig = {
        "orc_matrix_1q": (i, [vp, vp]),
        "orc_matrix_2q": (i, [vp, vp]),
        "orc_apply_1q_f64": (None, [vp, u32, u32, vp]),
        "orc_apply_2q_f64": (None, [vp, u32, u32, u32, vp]),
        "orc_apply_1q_f32": (None, [vp, u32, u32, vp]),
        "orc_apply_2q_f32": (None, [vp, u32, u32, u32, vp]),
        "orc_state_check": (i, [u32, i, i]),
        "orc_apply_gate": (i, [vp, i, u32, vp, vp, sz, C.POINTER(sz), vp, C.POINTER(sz)]),
        "orc_run_circuit": (i, [vp, i, u32, vp, sz, vp, sz, vp, C.POINTER(sz), C.POINTER(sz)]),
        "orc_run_circuit_noisy": (i, [vp, i, u32, vp, sz, d, d, d, vp, sz, vp, C.POINTER(sz), C.POINTER(sz),
                                      C.POINTER(sz)]),
        "orc_measure": (i, [vp, i, u32, u32, d, C.POINTER(i), C.POINTER(d)]),
        "orc_grover_optimal_iterations": (u32, [u32, sz]),
        "orc_run_grover": (i, [u32, vp, sz, C.c_int64, vp, vp, C.POINTER(C.c_uint64), C.POINTER(i),
                               C.POINTER(d), C.POINTER(u32)]),
        "orc_measure_all": (i, [vp, i, u32, vp, vp, vp]),
        "orc_run_shots": (i, [i, u32, vp, sz, sz, vp, sz, vp]),
        "orc_measure_all_shots": (i, [vp, i, u32, vp, sz, vp]),
        "orc_probabilities": (None, [vp, i, u32, vp]),
        "orc_probability_of_qubit": (d, [vp, u32, u32]),
        "orc_expectation": (d, [vp, u32, vp, vp, sz]),
        "orc_expectation_hamiltonian": (d, [vp, u32, vp, vp, vp, vp, sz]),
        "orc_normalize": (None, [vp, u32]),
        "orc_fidelity": (d, [vp, vp, u32]),
        "orc_circuit_depth": (u32, [u32, vp, sz]),
        "orc_mat_mul_2x2": (None, [vp, vp, vp]),
        "orc_fuse_gates": (sz, [vp, sz, vp]),
        "orc_build_qaoa_circuit": (sz, [u32, vp, vp, vp, sz, vp, vp, u32, vp]),
        "orc_cut_value": (d, [vp, vp, vp, sz, vp]),
        "orc_qaoa_cost": (i, [u32, vp, vp, vp, sz, vp, vp, u32, C.POINTER(d), vp]),
        "orc_run_qaoa": (i, [u32, vp, vp, vp, sz, u32, u32, d, vp, vp, vp, C.POINTER(QaoaResult)]),
        "orc_build_ansatz": (sz, [u32, u32, vp, vp]),
        "orc_energy": (i, [u32, vp, sz, vp, vp, vp, vp, sz, C.POINTER(d)]),
        "orc_run_vqe": (i, [u32, u32, vp, vp, vp, vp, sz, u32, d, d, vp, vp, C.POINTER(VqeResult)]),
        "orc_simd_apply_1q_best": (None, [vp, u32, u32, vp, i]),
        "orc_simd_apply_2q_best": (None, [vp, u32, u32, u32, vp, i]),
        "orc_simd_apply_1q_avx2": (None, [vp, u32, u32, vp]),
    }
    for name, (res, args) in sig.items():
        f = getattr(handle, name)
        f.restype = res
        f.argtypes = args
    return handle


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = load(LIB_PATH)
    return _lib


@contextlib.contextmanager
def planted(path):
    """Route every wrapper in this module through another oracle build (a
    mutant compiled with -DORC_PLANT=k by the pinning tests) while the block
    runs; the faithful build is restored afterwards."""
    global _lib
    saved = lib()
    _lib = load(path)
    try:
        yield
    finally:
        _lib = saved


def build_planted(k, out_dir):
    """Compile the scalar restatement with ORC_PLANT=k (one deliberate
    mistake, listed at the top of ruqu_oracle.cpp) plus the simd restatement
    into out_dir/liboracle_plant<k>.so; returns the path."""
    out = os.path.join(out_dir, "liboracle_plant%d.so" % k)
    flags = ["-std=c++17", "-O2", "-ffp-contract=off", "-fno-fast-math", "-fPIC", "-shared", "-pthread"]
    subprocess.run(["g++"] + flags + ["-DORC_PLANT=%d" % k, os.path.join(HERE, "ruqu_oracle.cpp"),
                                      os.path.join(HERE, "_build", "simd.o"), "-o", out], check=True)
    return out


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def gate_array(gates):
    """Accepts an rq_gate-compatible numpy array (ruqu_amd's GATE_DTYPE has the
    same layout) and returns it as the oracle's dtype."""
    a = np.ascontiguousarray(gates)
    return a.view(GATE_DTYPE) if a.dtype != GATE_DTYPE else a


def matrix_1q(g):
    out = np.zeros(4, dtype=np.complex128)
    rc = lib().orc_matrix_1q(_p(gate_array(np.array([g]))), _p(out))
    return None if rc else out.reshape(2, 2)


def matrix_2q(g):
    out = np.zeros(16, dtype=np.complex128)
    rc = lib().orc_matrix_2q(_p(gate_array(np.array([g]))), _p(out))
    return None if rc else out.reshape(4, 4)


def run_circuit(n, gates, f32=False, draws=None):
    """Simulator::run_with_config (no noise) from |0..0>; returns
    (amplitudes, outcomes, gate_count)."""
    g = gate_array(gates)
    amps = np.zeros(1 << n, dtype=np.complex64 if f32 else np.complex128)
    d = np.ascontiguousarray(draws if draws is not None else np.zeros(0), dtype=np.float64)
    outs = np.zeros(max(len(g), 1), dtype=OUTCOME_DTYPE)
    nout, cnt = C.c_size_t(0), C.c_size_t(0)
    rc = lib().orc_run_circuit(_p(amps), 1 if f32 else 0, n, _p(g), len(g), _p(d), len(d), _p(outs), C.byref(nout),
                               C.byref(cnt))
    if rc:
        raise RuntimeError("oracle run_circuit failed with status %d" % rc)
    return amps, outs[: nout.value].copy(), cnt.value


def run_circuit_noisy(n, gates, depol, bit_flip, phase_flip, draws, f32=False):
    """Simulator::run_with_config with a NoiseModel (simulator.rs:56-101,
    apply_noise :178-221) from |0..0>; returns (amplitudes, outcomes,
    gate_count, draws_used)."""
    g = gate_array(gates)
    amps = np.zeros(1 << n, dtype=np.complex64 if f32 else np.complex128)
    d = np.ascontiguousarray(draws, dtype=np.float64)
    outs = np.zeros(max(len(g), 1), dtype=OUTCOME_DTYPE)
    nout, cnt, used = C.c_size_t(0), C.c_size_t(0), C.c_size_t(0)
    rc = lib().orc_run_circuit_noisy(_p(amps), 1 if f32 else 0, n, _p(g), len(g), float(depol), float(bit_flip),
                                     float(phase_flip), _p(d), len(d), _p(outs), C.byref(nout), C.byref(cnt),
                                     C.byref(used))
    if rc:
        raise RuntimeError("oracle run_circuit_noisy failed with status %d" % rc)
    return amps, outs[: nout.value].copy(), cnt.value, used.value


def grover_optimal_iterations(n, m):
    return int(lib().orc_grover_optimal_iterations(n, m))


def run_grover(n, targets, iterations, draws):
    """grover.rs:98-183 run_grover (f64); iterations=None selects the optimal
    count.  Returns (final amplitudes, measured_state, target_found,
    success_probability, num_iterations)."""
    t = np.ascontiguousarray(targets, dtype=np.uint64)
    dr = np.ascontiguousarray(draws, dtype=np.float64)
    amps = np.zeros(1 << n, dtype=np.complex128)
    m, f, p, it = C.c_uint64(0), C.c_int(0), C.c_double(0), C.c_uint32(0)
    rc = lib().orc_run_grover(n, _p(t) if t.size else None, t.size, -1 if iterations is None else int(iterations),
                              _p(dr) if dr.size else None, _p(amps), C.byref(m), C.byref(f), C.byref(p), C.byref(it))
    if rc:
        raise RuntimeError("oracle run_grover failed with status %d" % rc)
    return amps, m.value, bool(f.value), p.value, it.value


def apply_gate(amps, n, gate, draw=0.0):
    g = gate_array(np.array([gate]))
    d = np.array([draw], dtype=np.float64)
    cur, nout = C.c_size_t(0), C.c_size_t(0)
    outs = np.zeros(1, dtype=OUTCOME_DTYPE)
    f32 = amps.dtype == np.complex64
    rc = lib().orc_apply_gate(_p(amps), 1 if f32 else 0, n, _p(g), _p(d), 1, C.byref(cur), _p(outs), C.byref(nout))
    return rc, (outs[0] if nout.value else None)


def measure(amps, n, q, draw):
    r, p = C.c_int(0), C.c_double(0)
    f32 = amps.dtype == np.complex64
    rc = lib().orc_measure(_p(amps), 1 if f32 else 0, n, q, draw, C.byref(r), C.byref(p))
    if rc:
        raise RuntimeError("oracle measure failed")
    return bool(r.value), p.value


def measure_all(amps, n, draws):
    res = np.zeros(n, dtype=np.int32)
    probs = np.zeros(n, dtype=np.float64)
    f32 = amps.dtype == np.complex64
    d = np.ascontiguousarray(draws, dtype=np.float64)
    lib().orc_measure_all(_p(amps), 1 if f32 else 0, n, _p(d), _p(res), _p(probs))
    return res, probs


def measure_all_shots(amps, n, draws):
    """measure_all (state.rs:329-335) on a fresh copy of `amps` for every row
    of draws (shots x n); returns the bits (shots x n, uint8).  Memoised over
    shared outcome prefixes, so 10^5 shots of a 26-qubit state take seconds."""
    a = np.ascontiguousarray(amps)
    d = np.ascontiguousarray(draws, dtype=np.float64)
    shots = d.shape[0]
    bits = np.zeros((shots, n), dtype=np.uint8)
    lib().orc_measure_all_shots(_p(a), 1 if a.dtype == np.complex64 else 0, n, _p(d), shots, _p(bits))
    return bits


def run_shots(n, gates, shots, draws, f32=False):
    g = gate_array(gates)
    d = np.ascontiguousarray(draws, dtype=np.float64).reshape(shots, -1)
    bits = np.zeros((shots, n), dtype=np.uint8)
    rc = lib().orc_run_shots(1 if f32 else 0, n, _p(g), len(g), shots, _p(d), d.shape[1], _p(bits))
    if rc:
        raise RuntimeError("oracle run_shots failed")
    return bits


def probabilities(amps, n):
    out = np.zeros(1 << n, dtype=np.float64)
    lib().orc_probabilities(_p(amps), 1 if amps.dtype == np.complex64 else 0, n, _p(out))
    return out


def probability_of_qubit(amps, n, q):
    return lib().orc_probability_of_qubit(_p(np.ascontiguousarray(amps, dtype=np.complex128)), n, q)


def expectation(amps, n, ops):
    """ops: [(qubit, op)] with op 0=I 1=X 2=Y 3=Z."""
    qs = np.array([q for q, _ in ops], dtype=np.uint32)
    os_ = np.array([o for _, o in ops], dtype=np.int32)
    return lib().orc_expectation(_p(np.ascontiguousarray(amps, dtype=np.complex128)), n, _p(qs), _p(os_), len(ops))


def expectation_hamiltonian(amps, n, terms):
    coeffs = np.array([c for c, _ in terms], dtype=np.float64)
    off = np.zeros(len(terms) + 1, dtype=np.uint64)
    qs, os_ = [], []
    for k, (_, ops) in enumerate(terms):
        off[k + 1] = off[k] + len(ops)
        qs += [q for q, _ in ops]
        os_ += [o for _, o in ops]
    qs = np.array(qs, dtype=np.uint32)
    os_ = np.array(os_, dtype=np.int32)
    return lib().orc_expectation_hamiltonian(_p(np.ascontiguousarray(amps, dtype=np.complex128)), n, _p(coeffs),
                                             _p(off), _p(qs), _p(os_), len(terms))


def normalize(amps, n):
    lib().orc_normalize(_p(amps), n)


def fidelity(a, b, n):
    return lib().orc_fidelity(_p(np.ascontiguousarray(a, dtype=np.complex128)),
                              _p(np.ascontiguousarray(b, dtype=np.complex128)), n)


def circuit_depth(n, gates):
    g = gate_array(gates)
    return lib().orc_circuit_depth(n, _p(g), len(g))


def fuse_gates(gates):
    g = gate_array(gates)
    out = np.zeros(max(len(g), 1), dtype=GATE_DTYPE)
    k = lib().orc_fuse_gates(_p(g), len(g), _p(out))
    return out[:k].copy()


def _graph_arrays(edges, weights):
    ei = np.array([e[0] for e in edges], dtype=np.uint32)
    ej = np.array([e[1] for e in edges], dtype=np.uint32)
    w = np.array(weights, dtype=np.float64)
    return ei, ej, w


def qaoa_circuit(n, edges, weights, gammas, betas):
    ei, ej, w = _graph_arrays(edges, weights)
    p = len(gammas)
    out = np.zeros(n + p * (len(edges) + n), dtype=GATE_DTYPE)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    k = lib().orc_build_qaoa_circuit(n, _p(ei), _p(ej), _p(w), len(edges), _p(g), _p(b), p, _p(out))
    return out[:k]


def qaoa_cost(n, edges, weights, gammas, betas, want_state=False):
    ei, ej, w = _graph_arrays(edges, weights)
    g = np.ascontiguousarray(gammas, dtype=np.float64)
    b = np.ascontiguousarray(betas, dtype=np.float64)
    c = C.c_double(0)
    amps = np.zeros(1 << n, dtype=np.complex128) if want_state else None
    rc = lib().orc_qaoa_cost(n, _p(ei), _p(ej), _p(w), len(edges), _p(g), _p(b), len(g), C.byref(c), _p(amps))
    if rc:
        raise RuntimeError("oracle qaoa_cost failed")
    return (c.value, amps) if want_state else c.value


def cut_value(edges, weights, bits):
    ei, ej, w = _graph_arrays(edges, weights)
    b = np.ascontiguousarray(bits, dtype=np.uint8)
    return lib().orc_cut_value(_p(ei), _p(ej), _p(w), len(edges), _p(b))


def run_qaoa(n, edges, weights, p, max_iterations, learning_rate):
    ei, ej, w = _graph_arrays(edges, weights)
    g = np.zeros(p)
    b = np.zeros(p)
    hist = np.zeros(max(max_iterations, 1))
    res = QaoaResult()
    rc = lib().orc_run_qaoa(n, _p(ei), _p(ej), _p(w), len(edges), p, max_iterations, learning_rate, _p(g), _p(b),
                            _p(hist), C.byref(res))
    if rc:
        raise RuntimeError("oracle run_qaoa failed")
    bits = [(res.best_index >> q) & 1 == 1 for q in range(n)]
    return {"best_cut_value": res.best_cut_value, "best_bitstring": bits, "optimal_gammas": g, "optimal_betas": b,
            "energy_history": hist[: res.iterations].copy(), "converged": bool(res.converged),
            "best_index": int(res.best_index)}


def _ham_arrays(terms):
    coeffs = np.array([c for c, _ in terms], dtype=np.float64)
    off = np.zeros(len(terms) + 1, dtype=np.uint64)
    qs, os_ = [], []
    for k, (_, ops) in enumerate(terms):
        off[k + 1] = off[k] + len(ops)
        qs += [q for q, _ in ops]
        os_ += [o for _, o in ops]
    return coeffs, off, np.array(qs, dtype=np.uint32), np.array(os_, dtype=np.int32)


def energy(n, gates, terms):
    coeffs, off, qs, os_ = _ham_arrays(terms)
    g = gate_array(gates)
    e = C.c_double(0)
    rc = lib().orc_energy(n, _p(g), len(g), _p(coeffs), _p(off), _p(qs), _p(os_), len(terms), C.byref(e))
    if rc:
        raise RuntimeError("oracle energy failed")
    return e.value


def run_vqe(n, depth, terms, max_iterations, threshold, learning_rate):
    coeffs, off, qs, os_ = _ham_arrays(terms)
    npar = 2 * n * depth
    best = np.zeros(npar)
    hist = np.zeros(max(max_iterations, 1))
    res = VqeResult()
    rc = lib().orc_run_vqe(n, depth, _p(coeffs), _p(off), _p(qs), _p(os_), len(terms), max_iterations, threshold,
                           learning_rate, _p(best), _p(hist), C.byref(res))
    if rc:
        raise RuntimeError("oracle run_vqe failed")
    return {"optimal_energy": res.optimal_energy, "optimal_parameters": best,
            "energy_history": hist[: res.num_iterations].copy(), "num_iterations": res.num_iterations,
            "converged": bool(res.converged)}


def build_ansatz(n, depth, params):
    out = np.zeros(depth * 3 * n, dtype=GATE_DTYPE)
    pr = np.ascontiguousarray(params, dtype=np.float64)
    k = lib().orc_build_ansatz(n, depth, _p(pr), _p(out))
    return out[:k]


def simd_apply_1q_best(amps, n, q, m, threads):
    lib().orc_simd_apply_1q_best(_p(amps), n, q, _p(np.ascontiguousarray(m, dtype=np.complex128)), threads)


def simd_apply_2q_best(amps, n, q1, q2, m, threads):
    lib().orc_simd_apply_2q_best(_p(amps), n, q1, q2, _p(np.ascontiguousarray(m, dtype=np.complex128)), threads)


def simd_apply_1q_avx2(amps, n, q, m):
    lib().orc_simd_apply_1q_avx2(_p(amps), n, q, _p(np.ascontiguousarray(m, dtype=np.complex128)))


def apply_1q(amps, n, q, m):
    if amps.dtype == np.complex64:
        lib().orc_apply_1q_f32(_p(amps), n, q, _p(np.ascontiguousarray(m, dtype=np.complex64)))
    else:
        lib().orc_apply_1q_f64(_p(amps), n, q, _p(np.ascontiguousarray(m, dtype=np.complex128)))


def apply_2q(amps, n, q1, q2, m):
    if amps.dtype == np.complex64:
        lib().orc_apply_2q_f32(_p(amps), n, q1, q2, _p(np.ascontiguousarray(m, dtype=np.complex64)))
    else:
        lib().orc_apply_2q_f64(_p(amps), n, q1, q2, _p(np.ascontiguousarray(m, dtype=np.complex128)))
