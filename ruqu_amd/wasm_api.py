"""The JS-facing surface of the reference's WASM crate
(crates/ruqu-wasm/src/lib.rs:70-503) over this package: the same names,
arguments, limits, error messages and result shapes (plain dicts where the
WASM returns plain JS objects), with the simulation on the device.

Differences: a JS `null` / `undefined` seed is None here; measurement draws
may be passed explicitly (the WASM draws them from the seeded StdRng).
"""
import time

import numpy as np

from .gates import PauliString, QuantumCircuit
from .grover import GroverConfig, run_grover
from .qaoa import Graph, build_qaoa_circuit
from .state import QuantumState, SimConfig, Simulator

WASM_MAX_QUBITS = 25  # lib.rs:38


def _limit_error(n):
    return ValueError("Qubit limit exceeded: %d requested, max %d in WASM" % (n, WASM_MAX_QUBITS))


class WasmQuantumCircuit:
    """lib.rs:70-203: a circuit builder limited to 25 qubits."""

    def __init__(self, num_qubits):
        if num_qubits > WASM_MAX_QUBITS:
            raise _limit_error(num_qubits)
        self.inner = QuantumCircuit(num_qubits)

    def h(self, qubit): self.inner.h(qubit)
    def x(self, qubit): self.inner.x(qubit)
    def y(self, qubit): self.inner.y(qubit)
    def z(self, qubit): self.inner.z(qubit)
    def s(self, qubit): self.inner.s(qubit)
    def t(self, qubit): self.inner.t(qubit)
    def rx(self, qubit, angle): self.inner.rx(qubit, angle)
    def ry(self, qubit, angle): self.inner.ry(qubit, angle)
    def rz(self, qubit, angle): self.inner.rz(qubit, angle)
    def cnot(self, control, target): self.inner.cnot(control, target)
    def cz(self, q1, q2): self.inner.cz(q1, q2)
    def swap(self, q1, q2): self.inner.swap(q1, q2)
    def rzz(self, q1, q2, angle): self.inner.rzz(q1, q2, angle)
    def measure(self, qubit): self.inner.measure(qubit)
    def measure_all(self): self.inner.measure_all()
    def reset(self, qubit): self.inner.reset(qubit)
    def barrier(self): self.inner.barrier()

    @property
    def num_qubits(self):
        return self.inner.num_qubits()

    @property
    def gate_count(self):
        return self.inner.gate_count()

    @property
    def depth(self):
        return self.inner.depth()


def simulate(circuit, draws=None):
    """lib.rs:254-284: Simulator::run, then
    {probabilities, measurements: [{qubit, result, probability}], num_qubits,
    gate_count, execution_time_ms}."""
    res = Simulator.run_with_config(circuit.inner, SimConfig(draws=draws))
    return {
        "probabilities": res.state.probabilities().tolist(),
        "measurements": [{"qubit": int(m.qubit), "result": bool(m.result), "probability": float(m.probability)}
                         for m in res.measurements],
        "num_qubits": int(res.metrics.num_qubits),
        "gate_count": int(res.metrics.gate_count),
        "execution_time_ms": res.metrics.execution_time_ns / 1_000_000.0,
    }


def estimate_memory(num_qubits):  # lib.rs:286-289 (QuantumState::estimate_memory: 16 * 2^n)
    return QuantumState.estimate_memory(num_qubits)


def max_qubits():  # lib.rs:292-295
    return WASM_MAX_QUBITS


def grover_search(num_qubits, target_states, seed=None, draws=None):
    """lib.rs:325-388: {measured_state, target_found, success_probability,
    num_iterations}."""
    if num_qubits > WASM_MAX_QUBITS:
        raise _limit_error(num_qubits)
    if seed is not None:
        seed = int(float(seed))  # `as_f64() as u64`
    r = run_grover(GroverConfig(num_qubits, [int(t) for t in target_states], None, seed), draws=draws)
    r.state.close()
    return {"measured_state": r.measured_state, "target_found": r.target_found,
            "success_probability": r.success_probability, "num_iterations": r.num_iterations}


def qaoa_maxcut(num_nodes, edges_flat, p, gammas, betas, seed=None):
    """lib.rs:407-503: unit-weight MaxCut QAOA state at the given angles:
    {probabilities, expected_cut}, expected_cut = sum over the listed edges of
    0.5 * (1 - <Z_i Z_j>)."""
    if num_nodes > WASM_MAX_QUBITS:
        raise _limit_error(num_nodes)
    if len(gammas) != p:
        raise ValueError("gammas length mismatch: expected %d (p), got %d" % (p, len(gammas)))
    if len(betas) != p:
        raise ValueError("betas length mismatch: expected %d (p), got %d" % (p, len(betas)))
    if len(edges_flat) % 2 != 0:
        raise ValueError("edges_flat must contain an even number of elements (pairs of node indices)")
    g = Graph(num_nodes)
    pairs = [(int(edges_flat[k]), int(edges_flat[k + 1])) for k in range(0, len(edges_flat), 2)]
    for i, j in pairs:
        g.add_edge(i, j, 1.0)
    circuit = build_qaoa_circuit(g, list(gammas), list(betas))
    res = Simulator.run_with_config(circuit, SimConfig(seed=None if seed is None else int(float(seed))))
    probs = res.state.probabilities()
    expected = 0.0
    for i, j in pairs:
        expected += 0.5 * (1.0 - res.state.expectation_value(PauliString([(i, 3), (j, 3)])))
    res.state.close()
    return {"probabilities": probs.tolist(), "expected_cut": expected}
