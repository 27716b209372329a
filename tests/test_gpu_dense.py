"""Dense 5-qubit blocks on the matrix cores (rq_run_circuit fusion = 2)
against the oracle: complex-f32 amplitudes within 1e-5 (the f32 bar) of the
FP64 oracle, for every gate variant, blocks padded with unused qubits,
2q gates in either qubit order, mid-circuit measurement (fed draws) and the
random-circuit workload; below 10 qubits the mode falls back to the fused
tile schedule."""
import numpy as np
import pytest

from ruqu_amd import workloads
from ruqu_amd.gates import Gate, QuantumCircuit
from ruqu_amd.state import SimConfig, Simulator
from tests.helpers import max_abs, mixed_circuit, orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,layers,seed", [(10, 3, 1), (12, 2, 2), (14, 2, 3), (16, 1, 4)])
def test_dense_mixed_circuits(n, layers, seed):
    c = mixed_circuit(n, layers, seed)
    want = orc.run_circuit(n, c.to_array())[0]
    got = Simulator.run_with_config(c, SimConfig(precision="c32", fusion="dense"))
    m = got.metrics
    print("n=%d gates=%d passes=%d" % (n, m.gate_count, m.passes), flush=True)
    assert max_abs(got.state.state_vector(), want) <= 1e-5
    assert m.passes < m.gate_count


def test_dense_small_blocks_and_orders():
    c = QuantumCircuit(11)
    c.h(10).cnot(10, 0).rzz(3, 7, 0.7).add_gate(Gate.SWAP(9, 1)).cz(5, 2).ry(4, 1.1)
    c.rx(6, 0.3).cnot(6, 10).t(8).rz(0, 2.0)
    want = orc.run_circuit(11, c.to_array())[0]
    got = Simulator.run_with_config(c, SimConfig(precision="c32", fusion=2))
    assert max_abs(got.state.state_vector(), want) <= 1e-5


def test_dense_random_circuit_and_measurement():
    n = 18
    c = workloads.random_circuit(n, 4, 9)
    c.measure(3)
    c2 = workloads.random_circuit(n, 2, 10)
    for g in c2.gates():
        c.add_gate(g)
    draws = np.array([0.37])
    want_state, want_meas, _ = orc.run_circuit(n, c.to_array(), draws=draws)
    got = Simulator.run_with_config(c, SimConfig(precision="c32", fusion="dense", draws=draws))
    assert [(o.qubit, o.result) for o in got.measurements] == [(int(m["qubit"]), int(m["result"])) for m in want_meas]
    assert max_abs(got.state.state_vector(), want_state) <= 1e-5


def test_dense_small_state_falls_back():
    c = mixed_circuit(6, 2, 5)
    want = orc.run_circuit(6, c.to_array())[0]
    got = Simulator.run_with_config(c, SimConfig(precision="c32", fusion="dense"))
    assert max_abs(got.state.state_vector(), want) <= 1e-5
