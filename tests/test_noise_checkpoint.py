"""Oracle pinning for the noise channel (simulator.rs:178-221) and the host
side of StateCheckpoint (replay.rs:251-294).  CPU only.

The noise checks expand a circuit by the reference's rules, written here
independently of the oracle (per gate: its Measure/Reset draw first, then for
each qubit of Gate::qubits in order: depolarising = one draw, below the rate
a second draw picks X/Y/Z at 1/3 and 2/3; bit flip = one draw; phase flip =
one draw; a channel of rate 0 consumes nothing), run the expanded circuit
through the oracle's noiseless path and require bit-identical amplitudes.
Each listed rule changes the result of at least one case below if broken.
"""
import struct

import numpy as np
import pytest

from oracle import oracle as orc
from ruqu_amd import Gate, GateKind, NoiseModel, QuantumCircuit, StateCheckpoint
from ruqu_amd import workloads
from tests.helpers import mixed_circuit

TWO_Q = {GateKind.CNOT, GateKind.CZ, GateKind.SWAP, GateKind.Rzz}


def expand(circuit, depol, bit_flip, phase_flip, draws):
    """(expanded circuit, measurement draws, draws used)."""
    out = QuantumCircuit(circuit.num_qubits())
    md, c = [], 0
    for g in circuit.gates():
        k = GateKind(g.kind)
        if k in (GateKind.Measure, GateKind.Reset):
            md.append(draws[c])
            c += 1
        out.add_gate(g)
        qs = [] if k == GateKind.Barrier else ([g.q0, g.q1] if k in TWO_Q else [g.q0])
        for q in qs:
            if depol > 0:
                r = draws[c]
                c += 1
                if r < depol:
                    ch = draws[c]
                    c += 1
                    out.add_gate(Gate.X(q) if ch < 1 / 3 else (Gate.Y(q) if ch < 2 / 3 else Gate.Z(q)))
            if bit_flip > 0:
                r = draws[c]
                c += 1
                if r < bit_flip:
                    out.add_gate(Gate.X(q))
            if phase_flip > 0:
                r = draws[c]
                c += 1
                if r < phase_flip:
                    out.add_gate(Gate.Z(q))
    return out, np.array(md), c


def measured_circuit(n, seed):
    c = mixed_circuit(n, 3, seed)
    c.measure(1)
    c.reset(0)
    c.h(0)
    c.cnot(0, n - 1)
    c.barrier()
    c.measure(n - 1)
    return c


RATES = [(0.0, 0.0, 0.0), (0.3, 0.0, 0.0), (0.0, 0.4, 0.0), (0.0, 0.0, 0.5), (0.2, 0.3, 0.25), (1.0, 1.0, 1.0)]


@pytest.mark.parametrize("rates", RATES)
@pytest.mark.parametrize("f32", [False, True])
def test_noisy_oracle_equals_expanded_circuit(rates, f32):
    n = 5
    c = measured_circuit(n, 11)
    draws = workloads.uniforms(1234, 8 * len(c.gates()) + 8)
    amps, outs, cnt, used = orc.run_circuit_noisy(n, c.to_array(), *rates, draws, f32=f32)
    ex, md, used_want = expand(c, *rates, draws)
    want, wouts, _ = orc.run_circuit(n, ex.to_array(), f32=f32, draws=md)
    assert used == used_want
    assert np.array_equal(amps, want)
    assert [tuple(o) for o in outs] == [tuple(o) for o in wouts]
    assert cnt == sum(1 for g in c.gates() if GateKind(g.kind) not in (GateKind.Measure, GateKind.Reset,
                                                                       GateKind.Barrier))


def test_noise_rate_zero_consumes_only_measurement_draws():
    n = 4
    c = measured_circuit(n, 3)
    draws = workloads.uniforms(5, 16)
    amps, _, _, used = orc.run_circuit_noisy(n, c.to_array(), 0.0, 0.0, 0.0, draws)
    want, _, _ = orc.run_circuit(n, c.to_array(), draws=draws)
    assert used == 2 + 1  # two Measure and one Reset
    assert np.array_equal(amps, want)


def test_bit_flip_certain_is_x_after_every_touched_qubit():
    # Rx(0.7) on qubit 0, then CNOT(0, 1): the channel with rate 1 applies X to
    # qubit 0 after the Rx and X to qubits 0 and 1 after the CNOT.
    c = QuantumCircuit(2).rx(0, 0.7).cnot(0, 1)
    draws = np.full(8, 0.5)
    amps, _, _, used = orc.run_circuit_noisy(2, c.to_array(), 0.0, 1.0, 0.0, draws)
    assert used == 3
    a0 = np.array([np.cos(0.35), -1j * np.sin(0.35)])          # Rx(0.7)|0>
    a0 = a0[::-1]                                                # X
    psi = np.zeros(4, dtype=complex)
    psi[0], psi[1] = a0                                          # index = q1 q0
    psi[[1, 3]] = psi[[3, 1]]                                    # CNOT(0 -> 1)
    psi = psi[[3, 2, 1, 0]]                                      # X on both qubits
    assert np.max(np.abs(amps - psi)) < 1e-15


def test_depolarizing_choice_thresholds():
    # One H gate on one qubit, depolarising with rate 1: the choice draw picks
    # X below 1/3, Y below 2/3, Z above.
    c = QuantumCircuit(1).h(0)
    plus = np.array([1, 1]) / np.sqrt(2)
    for choice, want in [(0.1, plus), (0.5, 1j * np.array([-1, 1]) / np.sqrt(2)), (0.9, np.array([1, -1]) / np.sqrt(2))]:
        amps, _, _, used = orc.run_circuit_noisy(1, c.to_array(), 1.0, 0.0, 0.0, np.array([0.0, choice]))
        assert used == 2
        assert np.max(np.abs(amps - want)) < 1e-15, choice


def test_noise_draws_exhausted_is_an_error():
    c = QuantumCircuit(2).h(0).h(1)
    with pytest.raises(RuntimeError):
        orc.run_circuit_noisy(2, c.to_array(), 0.0, 0.5, 0.0, np.array([0.9]))


def test_checkpoint_capture_restore():  # replay.rs:461-479
    rng = np.random.default_rng(3)
    amps = rng.normal(size=8) + 1j * rng.normal(size=8)
    cp = StateCheckpoint.capture(amps)
    assert np.array_equal(cp.restore(), amps)


def test_checkpoint_size_bytes():  # replay.rs:481-487
    cp = StateCheckpoint.capture(np.zeros(8, dtype=complex))
    assert cp.size_bytes() == 8 * 16


def test_checkpoint_byte_layout():
    cp = StateCheckpoint.capture(np.array([complex(1.5, -2.25), complex(-0.0, 3.0)]))
    assert cp.data == struct.pack("<dddd", 1.5, -2.25, -0.0, 3.0)


def test_noise_model_defaults():
    nm = NoiseModel()
    assert (nm.depolarizing_rate, nm.bit_flip_rate, nm.phase_flip_rate) == (0.0, 0.0, 0.0)
