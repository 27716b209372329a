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


def test_dense_full_size_matches_exact():
    """The benchmark circuit at full size (30 qubits complex f32, depth 40) as
    dense 5-qubit blocks against the exact per-gate f32 path, through
    size-independent checks on the device (FP64): both norms and the
    normalised fidelity."""
    from ruqu_amd.state import SimConfig as SC
    n = 30
    c = workloads.random_circuit(n, 40, 2)
    print("dense full size n=%d" % n, flush=True)
    exact = Simulator.run_with_config(c, SC(precision="c32"))
    print("  exact path done", flush=True)
    dense = Simulator.run_with_config(c, SC(precision="c32", fusion="dense"))
    nd, ne = dense.state.norm_sq(), exact.state.norm_sq()
    inf = 1.0 - exact.state.fidelity(dense.state) / (nd * ne)
    print("  norms %.9f %.9f  1-F %.3e  blocks %d for %d gates" % (nd, ne, inf, dense.metrics.passes,
                                                                 dense.metrics.gate_count), flush=True)
    assert abs(nd - 1.0) < 1e-4 and abs(ne - 1.0) < 1e-4
    assert abs(inf) < 1e-9
    assert dense.metrics.gate_count == exact.metrics.gate_count
    dense.state.close()
    exact.state.close()
