"""Variational Quantum Eigensolver on the device: vqe.rs:22-289.

The hardware-efficient ansatz (Ry then Rz on every qubit, then the CNOT
chain 0->1->...->n-1, per layer), the energy <psi|H|psi> and the
parameter-shift gradient descent loop follow run_vqe (vqe.rs:148-216) step
for step: energy, history, best (strictly lower), convergence test from the
second iteration, gradient of every parameter from two pi/2-shifted
evaluations at the same parameters, simultaneous update.  Each evaluation
resets one device-resident state to |0...0> and runs the ansatz through the
exact per-gate kernels (bit-identical f64 amplitudes); the energy is the
FP64 device reduction of expectation_hamiltonian.
"""
import math

import numpy as np

from .gates import Hamiltonian, PauliOp, PauliString, QuantumCircuit
from .state import QuantumState, SimConfig, Simulator


class VqeConfig:
    """vqe.rs:25-43."""

    def __init__(self, hamiltonian, num_qubits, ansatz_depth, max_iterations, convergence_threshold, learning_rate,
                 seed=None):
        self.hamiltonian = hamiltonian
        self.num_qubits = int(num_qubits)
        self.ansatz_depth = int(ansatz_depth)
        self.max_iterations = int(max_iterations)
        self.convergence_threshold = float(convergence_threshold)
        self.learning_rate = float(learning_rate)
        self.seed = seed


class VqeResult:
    """vqe.rs:46-57."""

    def __init__(self, optimal_energy, optimal_parameters, energy_history, num_iterations, converged):
        self.optimal_energy = optimal_energy
        self.optimal_parameters = optimal_parameters
        self.energy_history = energy_history
        self.num_iterations = num_iterations
        self.converged = converged


def num_parameters(num_qubits, depth):
    """vqe.rs:64-67: 2 * num_qubits * depth."""
    return 2 * int(num_qubits) * int(depth)


def build_ansatz(num_qubits, depth, params):
    """vqe.rs:80-111.  A wrong parameter count raises ValueError where the
    reference panics (assert_eq!, vqe.rs:82-88)."""
    expected = num_parameters(num_qubits, depth)
    if len(params) != expected:
        raise ValueError("build_ansatz: expected %d parameters, got %d" % (expected, len(params)))
    c = QuantumCircuit(num_qubits)
    idx = 0
    for _ in range(depth):
        for q in range(num_qubits):
            c.ry(q, float(params[idx]))
            idx += 1
        for q in range(num_qubits):
            c.rz(q, float(params[idx]))
            idx += 1
        for q in range(max(num_qubits - 1, 0)):
            c.cnot(q, q + 1)
    return c


class _Evaluator:
    """One device state reused for every evaluation of a run."""

    def __init__(self, config, precision):
        self.config = config
        self.precision = precision
        self.state = None

    def __call__(self, params):
        cfg = self.config
        circ = build_ansatz(cfg.num_qubits, cfg.ansatz_depth, params)
        if self.state is None:
            self.state = (QuantumState.new_with_seed(cfg.num_qubits, cfg.seed, self.precision)
                          if cfg.seed is not None else QuantumState.new(cfg.num_qubits, self.precision))
        # rq_run_circuit starts from |0...0> on the reused state
        Simulator.run_with_config(circ, SimConfig(seed=cfg.seed, precision=self.precision), state=self.state)
        return self.state.expectation_hamiltonian(cfg.hamiltonian)


def evaluate_energy(config, params, precision="c64"):
    """vqe.rs:120-129."""
    return _Evaluator(config, precision)(params)


def run_vqe(config, precision="c64"):
    """vqe.rs:148-216."""
    n_params = num_parameters(config.num_qubits, config.ansatz_depth)
    params = np.full(n_params, 0.1)
    history = []
    converged = False
    best_energy = 1.7976931348623157e308  # f64::MAX
    best_params = params.copy()
    ev = _Evaluator(config, precision)
    shift = math.pi / 2
    for iteration in range(config.max_iterations):
        energy = ev(params)
        history.append(energy)
        if energy < best_energy:
            best_energy = energy
            best_params = params.copy()
        if iteration > 0 and abs(history[iteration - 1] - energy) < config.convergence_threshold:
            converged = True
            break
        grad = np.zeros(n_params)
        for i in range(n_params):
            pp, pm = params.copy(), params.copy()
            pp[i] += shift
            pm[i] -= shift
            grad[i] = (ev(pp) - ev(pm)) / 2.0
        for i in range(n_params):
            params[i] -= config.learning_rate * grad[i]
    return VqeResult(best_energy, best_params, np.array(history), len(history), converged)


def h2_hamiltonian():
    """vqe.rs:231-267 (2-qubit H2 model; its exact ground energy is
    -1.857202, the smallest eigenvalue of the matrix it defines)."""
    return Hamiltonian([
        (-1.0523, PauliString([])),
        (0.3979, PauliString([(1, PauliOp.Z)])),
        (-0.3979, PauliString([(0, PauliOp.Z)])),
        (-0.0112, PauliString([(0, PauliOp.Z), (1, PauliOp.Z)])),
        (0.1809, PauliString([(0, PauliOp.X), (1, PauliOp.X)])),
    ], 2)


def single_z_hamiltonian():
    """vqe.rs:272-282: H = -Z."""
    return Hamiltonian([(-1.0, PauliString([(0, PauliOp.Z)]))], 1)
