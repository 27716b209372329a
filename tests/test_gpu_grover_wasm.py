"""Device Grover search and the WASM-shaped API against the oracle.

Bars: Grover's success probability, iteration count and measured state
equal the oracle's (the amplitudes before the measurement are produced by
the exact kernels, bit-identical); the final collapsed state within 1e-12;
sign-flip helpers bit-exact; WASM `simulate` probabilities bit-identical to
the oracle's |a|^2 and `qaoa_maxcut`'s expected cut within 1e-12.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from ruqu_amd import QuantumState, workloads
from ruqu_amd import wasm_api
from ruqu_amd.grover import GroverConfig, negate_amplitudes, negate_amplitudes_from, optimal_iterations, run_grover
from ruqu_amd.qaoa import Graph, build_qaoa_circuit
from tests.helpers import max_abs, random_state

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("n,targets,it", [(4, [7], None), (3, [5], None), (3, [0], None), (3, [1, 5], None),
                                          (3, [0, 2, 4, 6], None), (3, [3], 2), (10, [700, 3], None),
                                          (12, [4095], None), (5, [6, 6, 9], 3)])
def test_grover_matches_oracle(n, targets, it):
    flush("grover n=%d targets=%s it=%s" % (n, targets, it))
    draws = workloads.uniforms(n * 31 + len(targets), n)
    r = run_grover(GroverConfig(n, targets, it, 42), draws=draws)
    amps, m, found, p, iters = orc.run_grover(n, targets, it, draws)
    assert r.num_iterations == iters
    assert r.success_probability == p
    assert r.measured_state == m and r.target_found == found
    assert max_abs(r.state.state_vector(), amps) < 1e-12
    r.state.close()


def test_optimal_iterations_device_equals_oracle():
    for n in range(1, 21):
        for m in (1, 2, 3, 5, 1 << max(0, n - 1)):
            assert optimal_iterations(n, m) == orc.grover_optimal_iterations(n, m), (n, m)


@pytest.mark.parametrize("prec", ["c64", "c32"])
def test_negate_helpers_bit_exact(prec):
    n = 10
    psi = random_state(n, 4, np.complex128 if prec == "c64" else np.complex64)
    st = QuantumState.from_amplitudes(psi, n, prec)
    negate_amplitudes(st, [3, 700, 3, 12])
    want = psi.copy()
    want[[700, 12]] *= -1
    assert np.array_equal(st.state_vector(), want)
    negate_amplitudes_from(st, 1)
    want[1:] *= -1
    assert np.array_equal(st.state_vector(), want)


def test_wasm_simulate_shape_and_values():
    qc = wasm_api.WasmQuantumCircuit(6)
    qc.h(0)
    qc.cnot(0, 1)
    qc.rz(2, 0.4)
    qc.ry(3, 1.1)
    qc.rzz(3, 5, 0.7)
    qc.measure(1)
    draws = np.array([0.25])
    out = wasm_api.simulate(qc, draws=draws)
    assert set(out) == {"probabilities", "measurements", "num_qubits", "gate_count", "execution_time_ms"}
    want, wouts, cnt = orc.run_circuit(6, qc.inner.to_array(), draws=draws)
    assert out["probabilities"] == list(want.real ** 2 + want.imag ** 2)
    assert out["num_qubits"] == 6 and out["gate_count"] == cnt == 5
    assert [(m["qubit"], m["result"]) for m in out["measurements"]] == [(int(o[0]), bool(o[1])) for o in wouts]
    assert qc.num_qubits == 6 and qc.gate_count == 6 and qc.depth >= 2
    assert wasm_api.estimate_memory(10) == 16 << 10


def test_wasm_grover_and_qaoa():
    g = wasm_api.grover_search(4, [7], seed=42, draws=np.full(4, 0.5))
    assert set(g) == {"measured_state", "target_found", "success_probability", "num_iterations"}
    _, m, found, p, iters = orc.run_grover(4, [7], None, np.full(4, 0.5))
    assert (g["measured_state"], g["target_found"], g["success_probability"], g["num_iterations"]) == (m, found, p, iters)
    edges = [0, 1, 1, 2, 2, 3, 3, 0, 0, 2]
    q = wasm_api.qaoa_maxcut(4, edges, 2, [0.4, 0.8], [0.7, 0.3])
    assert set(q) == {"probabilities", "expected_cut"}
    gr = Graph(4)
    for i in range(0, len(edges), 2):
        gr.add_edge(edges[i], edges[i + 1], 1.0)
    amps, _, _ = orc.run_circuit(4, build_qaoa_circuit(gr, [0.4, 0.8], [0.7, 0.3]).to_array())
    assert q["probabilities"] == list(amps.real ** 2 + amps.imag ** 2)
    exp = sum(0.5 * (1 - orc.expectation(amps, 4, [(edges[k], 3), (edges[k + 1], 3)])) for k in range(0, 10, 2))
    assert abs(q["expected_cut"] - exp) < 1e-12
