"""Known-answer tests restated from the reference's own test files
(crates/ruqu-core/tests/test_state.rs, test_gates.rs, test_simulator.rs,
crates/ruqu-core/src/mixed_precision.rs:597-746,
crates/ruqu-algorithms/tests/test_algorithms.rs:19-175).

Each test runs on every backend: the oracle (CPU restatement, always) and
the device engine in c64, c32 and fused mode (marked gpu).  Tolerances are the
reference's: EPSILON = 1e-10 (test_state.rs:5) for f64, 1e-5 for f32
(mixed_precision.rs:669-690).
"""
import math

import numpy as np
import pytest

from ruqu_amd.gates import Gate, GateKind, QuantumCircuit
from tests.helpers import orc

S2 = 1.0 / math.sqrt(2.0)


class OracleBackend:
    name = "oracle"
    eps = 1e-10

    def __init__(self, f32=False):
        self.f32 = f32
        if f32:
            self.eps = 1e-5

    def run(self, n, gates, draws=None):
        amps, outs, _ = orc.run_circuit(n, QuantumCircuit.from_array(n, _arr(n, gates)).to_array(), self.f32, draws)
        return amps.astype(np.complex128), outs

    def expectation(self, n, gates, ops):
        amps, _ = self.run(n, gates)
        return orc.expectation(amps, n, ops)


class GpuBackend:
    def __init__(self, precision, fusion):
        self.precision = precision
        self.fusion = fusion
        self.name = "gpu-%s%s" % (precision, "-fused" if fusion else "")
        self.eps = 1e-10 if precision == "c64" else 1e-5

    def _run(self, n, gates, draws=None):
        from ruqu_amd import SimConfig, Simulator
        c = QuantumCircuit.from_array(n, _arr(n, gates))
        res = Simulator.run_with_config(c, SimConfig(seed=1, precision=self.precision, fusion=self.fusion,
                                                     draws=draws))
        return res

    def run(self, n, gates, draws=None):
        res = self._run(n, gates, draws)
        outs = np.array([(m.qubit, int(m.result), m.probability) for m in res.measurements],
                        dtype=orc.OUTCOME_DTYPE)
        return res.state.state_vector().astype(np.complex128), outs

    def expectation(self, n, gates, ops):
        from ruqu_amd import PauliString
        res = self._run(n, gates)
        return res.state.expectation_value(PauliString(ops))


def _arr(n, gates):
    c = QuantumCircuit(n)
    for g in gates:
        c.add_gate(g)
    return c.to_array()


BACKENDS = [
    pytest.param(OracleBackend(), id="oracle"),
    pytest.param(OracleBackend(f32=True), id="oracle-f32"),
    pytest.param(GpuBackend("c64", False), id="gpu-c64", marks=pytest.mark.gpu),
    pytest.param(GpuBackend("c32", False), id="gpu-c32", marks=pytest.mark.gpu),
    pytest.param(GpuBackend("c64", True), id="gpu-c64-fused", marks=pytest.mark.gpu),
    pytest.param(GpuBackend("c32", True), id="gpu-c32-fused", marks=pytest.mark.gpu),
]


def close(a, b, eps):
    return abs(a - b) < eps


@pytest.mark.parametrize("be", BACKENDS)
class TestStateKAT:
    def test_hadamard_amplitudes(self, be):  # test_state.rs:80-89
        sv, _ = be.run(1, [Gate.H(0)])
        assert close(sv[0].real, S2, be.eps) and close(sv[0].imag, 0, be.eps)
        assert close(sv[1].real, S2, be.eps) and close(sv[1].imag, 0, be.eps)

    def test_double_hadamard(self, be):  # :92-100
        sv, _ = be.run(1, [Gate.H(0), Gate.H(0)])
        assert close(abs(sv[0]) ** 2, 1.0, be.eps) and close(abs(sv[1]) ** 2, 0.0, be.eps)

    def test_hadamard_on_qubit_1_of_2(self, be):  # :103-113 (little-endian pin)
        sv, _ = be.run(2, [Gate.H(1)])
        p = np.abs(sv) ** 2
        assert np.allclose(p, [0.5, 0.0, 0.5, 0.0], atol=be.eps)

    def test_y_on_zero(self, be):  # :155-165: Y|0> = i|1>
        sv, _ = be.run(1, [Gate.Y(0)])
        assert close(sv[1].real, 0.0, be.eps) and close(sv[1].imag, 1.0, be.eps)

    def test_z_on_plus(self, be):  # :181-195
        sv, _ = be.run(1, [Gate.H(0), Gate.Z(0)])
        assert close(sv[0].real, S2, be.eps) and close(sv[1].real, -S2, be.eps)

    def test_bell_state(self, be):  # :214-230
        sv, _ = be.run(2, [Gate.H(0), Gate.CNOT(0, 1)])
        assert np.allclose(np.abs(sv) ** 2, [0.5, 0, 0, 0.5], atol=be.eps)

    def test_ghz(self, be):  # :280-307
        sv, _ = be.run(3, [Gate.H(0), Gate.CNOT(0, 1), Gate.CNOT(1, 2)])
        p = np.abs(sv) ** 2
        assert close(p[0], 0.5, be.eps) and close(p[7], 0.5, be.eps)

    def test_cnot_control_is_first(self, be):  # gate.rs:159-166
        sv, _ = be.run(2, [Gate.X(0), Gate.CNOT(0, 1)])
        assert close(abs(sv[3]) ** 2, 1.0, be.eps)
        sv, _ = be.run(2, [Gate.X(1), Gate.CNOT(0, 1)])
        assert close(abs(sv[2]) ** 2, 1.0, be.eps)

    def test_swap(self, be):  # :314-345
        sv, _ = be.run(2, [Gate.X(0), Gate.SWAP(0, 1)])
        assert close(abs(sv[2]) ** 2, 1.0, be.eps)

    def test_rx_pi_flips(self, be):  # :360-368
        sv, _ = be.run(1, [Gate.Rx(0, math.pi)])
        assert close(abs(sv[1]) ** 2, 1.0, be.eps)

    def test_ry_pi_flips(self, be):  # :371-378
        sv, _ = be.run(1, [Gate.Ry(0, math.pi)])
        assert close(abs(sv[1]) ** 2, 1.0, be.eps)

    def test_rz_preserves_probability(self, be):  # :381-386
        sv, _ = be.run(1, [Gate.Rz(0, 1.234)])
        assert close(abs(sv[0]) ** 2, 1.0, be.eps)

    def test_rx_half_pi(self, be):  # :389-398
        sv, _ = be.run(1, [Gate.Rx(0, math.pi / 2)])
        assert np.allclose(np.abs(sv) ** 2, [0.5, 0.5], atol=be.eps)

    def test_cz_on_11(self, be):  # :416-426
        sv, _ = be.run(2, [Gate.X(0), Gate.X(1), Gate.CZ(0, 1)])
        assert close(sv[3].real, -1.0, be.eps)

    def test_cz_on_01(self, be):  # :429-436
        sv, _ = be.run(2, [Gate.X(1), Gate.CZ(0, 1)])
        assert close(sv[2].real, 1.0, be.eps)

    def test_s_gate_on_plus(self, be):  # :880-891
        sv, _ = be.run(1, [Gate.H(0), Gate.S(0)])
        assert close(sv[0].real, S2, be.eps) and close(sv[1].real, 0.0, be.eps) and close(sv[1].imag, S2, be.eps)

    def test_t_gate_phase(self, be):  # :894-905
        sv, _ = be.run(1, [Gate.H(0), Gate.T(0)])
        ph = math.pi / 4
        assert close(sv[1].real, S2 * math.cos(ph), be.eps) and close(sv[1].imag, S2 * math.sin(ph), be.eps)

    def test_rzz_zero_is_identity(self, be):  # :912-922
        pre = [Gate.H(0), Gate.Ry(1, 0.7)]
        sv0, _ = be.run(2, pre)
        sv1, _ = be.run(2, pre + [Gate.Rzz(0, 1, 0.0)])
        assert np.allclose(np.abs(sv0) ** 2, np.abs(sv1) ** 2, atol=be.eps)

    @pytest.mark.parametrize("bits,idx", [((0, 0), 0), ((0, 1), 2), ((1, 0), 1), ((1, 1), 3)])
    def test_superdense_coding(self, be, bits, idx):  # :976-1029
        g = [Gate.H(0), Gate.CNOT(0, 1)]
        if bits[1]:
            g.append(Gate.X(0))
        if bits[0]:
            g.append(Gate.Z(0))
        g += [Gate.CNOT(0, 1), Gate.H(0)]
        sv, _ = be.run(2, g)
        assert close(abs(sv[idx]) ** 2, 1.0, be.eps)

    @pytest.mark.parametrize("ops,prep,want", [
        ([(0, 3)], [], 1.0),                                   # :556-562
        ([(0, 3)], [Gate.X(0)], -1.0),                         # :566-573
        ([(0, 3)], [Gate.H(0)], 0.0),                          # :577-584
        ([(0, 1)], [Gate.H(0)], 1.0),                          # :588-595
        ([(0, 1)], [Gate.X(0), Gate.H(0)], -1.0),              # :599-607
        ([(0, 3), (1, 3)], [Gate.H(0), Gate.CNOT(0, 1)], 1.0),  # :611-619
        ([(0, 1), (1, 1)], [Gate.H(0), Gate.CNOT(0, 1)], 1.0),  # :623-631
        ([(0, 2), (1, 2)], [Gate.H(0), Gate.CNOT(0, 1)], -1.0),  # :635-643
        ([], [Gate.H(0)], 1.0),                                # :647-653
    ])
    def test_pauli_expectations(self, be, ops, prep, want):
        n = 2 if any(q == 1 for q, _ in ops) or len(prep) > 1 else 1
        assert close(be.expectation(n, prep, ops), want, be.eps)

    def test_measure_deterministic_one(self, be):  # :465-471, mixed_precision.rs:704-712
        _, outs = be.run(1, [Gate.X(0), Gate.Measure(0)], draws=[0.3])
        assert outs[0]["result"] == 1 and close(outs[0]["probability"], 1.0, be.eps)

    def test_measure_collapses(self, be):  # :443-453
        sv, outs = be.run(1, [Gate.H(0), Gate.Measure(0)], draws=[0.75])
        assert outs[0]["result"] == 1 and close(abs(sv[1]) ** 2, 1.0, be.eps)
        sv, outs = be.run(1, [Gate.H(0), Gate.Measure(0)], draws=[0.25])
        assert outs[0]["result"] == 0 and close(abs(sv[0]) ** 2, 1.0, be.eps)

    def test_measure_bell_correlated(self, be):  # :474-482
        for d in (0.1, 0.9):
            _, outs = be.run(2, [Gate.H(0), Gate.CNOT(0, 1), Gate.Measure(0), Gate.Measure(1)], draws=[d, 0.5])
            assert outs[0]["result"] == outs[1]["result"]

    def test_reset_from_one(self, be):  # :764-770
        sv, _ = be.run(1, [Gate.X(0), Gate.Reset(0)], draws=[0.5])
        assert close(abs(sv[0]) ** 2, 1.0, be.eps)

    def test_reset_one_qubit_of_entangled(self, be):  # :782-791
        for d in (0.2, 0.8):
            sv, _ = be.run(2, [Gate.H(0), Gate.CNOT(0, 1), Gate.Reset(0)], draws=[d])
            p = np.abs(sv) ** 2
            assert close(p[1] + p[3], 0.0, be.eps)

    def test_teleportation(self, be):  # :929-969, with every measurement branch forced
        target = math.sin(1.23 / 2) ** 2
        for d0 in (0.1, 0.9):
            for d1 in (0.1, 0.9):
                g = [Gate.Ry(0, 1.23), Gate.H(1), Gate.CNOT(1, 2), Gate.CNOT(0, 1), Gate.H(0), Gate.Measure(0),
                     Gate.Measure(1)]
                sv, outs = be.run(3, g, draws=[d0, d1])
                m0, m1 = int(outs[0]["result"]), int(outs[1]["result"])
                fix = ([Gate.X(2)] if m1 else []) + ([Gate.Z(2)] if m0 else [])
                # apply the corrections on the collapsed state with the oracle
                amps = sv.copy()
                for gg in fix:
                    orc.apply_gate(amps, 3, QuantumCircuit(3).add_gate(gg).to_array()[0])
                p2 = float(np.sum(np.abs(amps[4:]) ** 2))
                assert abs(p2 - target) < 0.01

    @pytest.mark.parametrize("oracle_kind,balanced", [("f0", False), ("f1", False), ("f2", True), ("f3", True)])
    def test_deutsch(self, be, oracle_kind, balanced):  # test_algorithms.rs:19-120
        g = [Gate.X(1), Gate.H(0), Gate.H(1)]
        g += {"f0": [], "f1": [Gate.X(1)], "f2": [Gate.CNOT(0, 1)],
              "f3": [Gate.X(0), Gate.CNOT(0, 1), Gate.X(0)]}[oracle_kind]
        g.append(Gate.H(0))
        sv, _ = be.run(2, g)
        p = np.abs(sv) ** 2
        assert (p[1] + p[3] > 0.5) == balanced

    def test_normalization_many_gates(self, be):  # :745-757
        g = []
        for k in range(50):
            g += [Gate.H(k % 3), Gate.Rx((k + 1) % 3, 0.3 * k), Gate.CNOT(k % 3, (k + 1) % 3)]
        sv, _ = be.run(3, g)
        assert close(float(np.sum(np.abs(sv) ** 2)), 1.0, be.eps)


def test_gate_matrices_unitary_and_special_angles():  # test_gates.rs:22-87, :364-435
    for k in [GateKind.H, GateKind.X, GateKind.Y, GateKind.Z, GateKind.S, GateKind.Sdg, GateKind.T, GateKind.Tdg,
              GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.Phase]:
        for th in (0.0, 0.37, math.pi, 2 * math.pi):
            m = orc.matrix_1q(QuantumCircuit(1).add_gate(Gate(k, 0, 0, th)).to_array()[0])
            assert np.allclose(m @ m.conj().T, np.eye(2), atol=1e-12)
    rx0 = orc.matrix_1q(QuantumCircuit(1).rx(0, 0.0).to_array()[0])
    assert np.allclose(rx0, np.eye(2), atol=1e-15)
    rxpi = orc.matrix_1q(QuantumCircuit(1).rx(0, math.pi).to_array()[0])
    assert np.allclose(rxpi, [[0, -1j], [-1j, 0]], atol=1e-15)
    for k in [GateKind.CNOT, GateKind.CZ, GateKind.SWAP, GateKind.Rzz]:
        m = orc.matrix_2q(QuantumCircuit(2).add_gate(Gate(k, 0, 1, 0.9)).to_array()[0])
        assert np.allclose(m @ m.conj().T, np.eye(4), atol=1e-12)


def test_gate_algebra_identities():  # test_gates.rs:223 (XY=iZ), :282 (S^2=Z), :331 (T^2=S)
    def m(k):
        return orc.matrix_1q(QuantumCircuit(1).add_gate(Gate(k, 0)).to_array()[0])
    assert np.allclose(m(GateKind.X) @ m(GateKind.Y), 1j * m(GateKind.Z))
    assert np.allclose(m(GateKind.S) @ m(GateKind.S), m(GateKind.Z))
    assert np.allclose(m(GateKind.T) @ m(GateKind.T), m(GateKind.S))


def test_circuit_depth_kats():  # test_simulator.rs:110-138
    assert QuantumCircuit(1).h(0).depth() == 1
    assert QuantumCircuit(3).h(0).h(1).h(2).depth() == 1
    assert QuantumCircuit(1).h(0).x(0).y(0).depth() == 3
    assert QuantumCircuit(3).h(0).h(1).h(2).cnot(0, 1).cnot(1, 2).depth() == 3
    c = QuantumCircuit(3).h(0).barrier().h(1)
    assert c.depth() == 2
    for circ in (QuantumCircuit(3).h(0).h(1).h(2).cnot(0, 1).cnot(1, 2), QuantumCircuit(2).h(0).barrier().cnot(0, 1)):
        assert circ.depth() == orc.circuit_depth(circ.num_qubits(), circ.to_array())


def test_memory_estimate():  # test_state.rs:837-852, mixed_precision.rs:668-674
    from ruqu_amd import QuantumState
    assert QuantumState.estimate_memory(1) == 32
    assert QuantumState.estimate_memory(10) == 16384
    assert QuantumState.estimate_memory(20) == 16_777_216
    assert QuantumState.estimate_memory(3, "c32") == 64
    assert QuantumState.estimate_memory(10, "c32") == 8192
