"""Grover search (crates/ruqu-algorithms/src/grover.rs) on the device.

run_grover keeps the reference's sequence gate for gate (exact kernels, so
amplitudes are bit-identical to the reference's f64 loops): H on every
qubit, then per iteration the oracle sign flips it performs through
amplitudes_mut, H on every qubit, the sign flip of every amplitude but index
0, H on every qubit; then the targets' probabilities and the measurement of
every qubit.  The measurement draws (one uniform per qubit) are inputs, as
everywhere in this package.
"""
import ctypes as C

import numpy as np

from . import _native
from ._native import lib
from .rng import state_rng
from .state import QuantumState, _as_prec, check


class RqGroverResult(C.Structure):
    _fields_ = [("measured_state", C.c_uint64), ("target_found", C.c_int32), ("num_iterations", C.c_uint32),
                ("success_probability", C.c_double)]


_native.register("rq_grover_optimal_iterations", C.c_uint32, [C.c_uint32, C.c_size_t])
_native.register("rq_grover_run", C.c_int, [C.c_uint32, C.c_void_p, C.c_size_t, C.c_int64, C.c_int, C.c_int,
                                            C.c_void_p, C.POINTER(RqGroverResult), C.POINTER(C.c_void_p)])
_native.register("rq_negate_amplitudes", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t])
_native.register("rq_negate_amplitudes_from", C.c_int, [C.c_void_p, C.c_uint64])


class GroverConfig:  # grover.rs:28-40
    def __init__(self, num_qubits, target_states, num_iterations=None, seed=None):
        self.num_qubits = int(num_qubits)
        self.target_states = [int(t) for t in target_states]
        self.num_iterations = num_iterations
        self.seed = seed


class GroverResult:  # grover.rs:42-55
    def __init__(self, measured_state, target_found, success_probability, num_iterations, state):
        self.measured_state = measured_state
        self.target_found = target_found
        self.success_probability = success_probability
        self.num_iterations = num_iterations
        self.state = state


def optimal_iterations(num_qubits, num_targets):  # grover.rs:69-74
    return int(lib().rq_grover_optimal_iterations(int(num_qubits), int(num_targets)))


def run_grover(config, precision="c64", draws=None):  # grover.rs:98-183
    """draws: num_qubits uniforms for the final measurement (default: the
    stream of StdRng seeded with config.seed, as the reference's state)."""
    n = config.num_qubits
    t = np.ascontiguousarray(config.target_states, dtype=np.uint64)
    if draws is None:  # grover.rs:118-120: the state's StdRng, or entropy without a seed
        draws = state_rng(config.seed).random(n)
    d = np.ascontiguousarray(draws, dtype=np.float64)
    if d.size < n:
        raise ValueError("run_grover needs %d measurement draws, got %d" % (n, d.size))
    it = -1 if config.num_iterations is None else int(config.num_iterations)
    p = _as_prec(precision)
    res = RqGroverResult()
    h = C.c_void_p()
    check(lib().rq_grover_run(n, t.ctypes.data if t.size else None, t.size, it, p,
                              0 if config.seed is None else 1, d.ctypes.data if d.size else None, C.byref(res),
                              C.byref(h)))
    return GroverResult(int(res.measured_state), bool(res.target_found), float(res.success_probability),
                        int(res.num_iterations), QuantumState(h.value, p, config.seed))


def negate_amplitudes(state, indices):
    """a[i] = -a[i] for every listed index (the oracle step of grover.rs:130-139)."""
    idx = np.ascontiguousarray(indices, dtype=np.uint64)
    check(lib().rq_negate_amplitudes(state.handle, idx.ctypes.data if idx.size else None, idx.size))


def negate_amplitudes_from(state, start):
    """a[i] = -a[i] for every i >= start (grover.rs:146-155 with start = 1)."""
    check(lib().rq_negate_amplitudes_from(state.handle, int(start)))
