"""Seeded runs on the device reproduce the reference's outcomes: the draws
come from StdRng::seed_from_u64(seed) (state.rs:68), and the oracle, given
the same stream from its own restatement (oracle/stdrng_oracle.py), makes
the same decisions.  Bars: outcomes identical, amplitudes within 1e-12
(f64, exact kernels)."""
import numpy as np
import pytest

from oracle import oracle as orc
from oracle.stdrng_oracle import StdRngOracle
from ruqu_amd import NoiseModel, QuantumCircuit, QuantumState, SimConfig, Simulator
from ruqu_amd.grover import GroverConfig, run_grover
from tests.helpers import max_abs

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


@pytest.mark.parametrize("seed", [0, 42, 12345, 99])
def test_new_with_seed_measure(seed):
    flush("new_with_seed measure seed=%d" % seed)
    st = QuantumState.new_with_seed(3, seed)
    c = QuantumCircuit(3).h(0).ry(1, 1.23).cnot(1, 2).h(2)
    for g in c.gates():
        st.apply_gate(g)
    outs = [st.measure(q) for q in range(3)]
    draws = StdRngOracle(seed).random(3)
    amps, _, _ = orc.run_circuit(3, c.to_array())
    for q in range(3):
        r, p = orc.measure(amps, 3, q, draws[q])
        assert outs[q].result == r and abs(outs[q].probability - p) < 1e-12
    assert max_abs(st.state_vector(), amps) < 1e-12
    st.close()


@pytest.mark.parametrize("seed", [3, 7])
def test_run_with_config_seeded(seed):
    flush("run_with_config seed=%d" % seed)
    c = QuantumCircuit(4).h(0).cnot(0, 1).measure(1).rx(2, 0.7).cnot(2, 3).measure(3).reset(0).h(0).measure(0)
    res = Simulator.run_with_config(c, SimConfig(seed=seed))
    amps, outs, _ = orc.run_circuit(4, c.to_array(), draws=StdRngOracle(seed).random(8))
    assert [(m.qubit, m.result) for m in res.measurements] == [(int(o["qubit"]), bool(o["result"])) for o in outs]
    assert max_abs(res.state.state_vector(), amps) < 1e-12


def test_run_with_config_seeded_noise():
    flush("noisy run seed=11")
    c = QuantumCircuit(3).h(0).cnot(0, 1).rz(2, 0.3).cnot(1, 2).measure(2)
    nm = NoiseModel(depolarizing_rate=0.2, bit_flip_rate=0.1, phase_flip_rate=0.1)
    res = Simulator.run_with_config(c, SimConfig(seed=11, noise=nm))
    amps, outs, _, _ = orc.run_circuit_noisy(3, c.to_array(), 0.2, 0.1, 0.1, StdRngOracle(11).random(64))
    assert [(m.qubit, m.result) for m in res.measurements] == [(int(o["qubit"]), bool(o["result"])) for o in outs]
    assert max_abs(res.state.state_vector(), amps) < 1e-12


def test_run_shots_seeded_counts():
    flush("run_shots seed=5")
    c = QuantumCircuit(3).h(0).cnot(0, 1).ry(2, 0.9)
    r = Simulator.run_shots(c, 200, seed=5)
    amps, _, _ = orc.run_circuit(3, c.to_array())
    want = {}
    for s in range(200):
        res, _ = orc.measure_all(amps.copy(), 3, StdRngOracle(5 + s).random(3))
        key = tuple(bool(b) for b in res)
        want[key] = want.get(key, 0) + 1
    assert r.counts == want


@pytest.mark.parametrize("n,targets,seed", [(4, [7], 42), (6, [5, 33], 9)])
def test_grover_seeded(n, targets, seed):
    flush("grover seed=%d" % seed)
    r = run_grover(GroverConfig(n, targets, None, seed))
    _, m, found, p, it = orc.run_grover(n, targets, None, StdRngOracle(seed).random(n))
    assert (r.measured_state, r.target_found, r.num_iterations) == (m, found, it)
    assert r.success_probability == p
    r.state.close()
