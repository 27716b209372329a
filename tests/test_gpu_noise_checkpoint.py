"""Device noise trajectories (simulator.rs:178-221) and StateCheckpoint
(replay.rs:251-294) against the oracle.

Bars: given the same draws, a noisy run of unitary gates through the exact
kernels is bit-identical to the oracle's (amplitudes, draws consumed, gate
count); with Measure / Reset gates the outcomes and draws consumed are
identical and amplitudes and probabilities are within 1e-12 (f64) / 1e-5 (f32)
(the collapse probability is a device reduction in another order; the
reference's f32 state sums it in f32); the fused schedule is within the same
bars; checkpoint bytes equal StateCheckpoint::capture of the amplitudes and
restores are exact.
"""
import numpy as np
import pytest

from oracle import oracle as orc
from ruqu_amd import InvalidStateVector, NoiseModel, QuantumState, SimConfig, Simulator, StateCheckpoint
from ruqu_amd import workloads
from tests.helpers import max_abs, mixed_circuit

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


def noisy_circuit(n, seed):
    c = mixed_circuit(n, 4, seed)
    c.measure(2)
    c.reset(1)
    c.h(1)
    c.cnot(1, n - 1)
    c.barrier()
    c.measure(0)
    return c


RATES = [(0.0, 0.3, 0.0), (0.25, 0.0, 0.0), (0.1, 0.2, 0.3), (1.0, 1.0, 1.0)]
TOL = {"c64": 1e-12, "c32": 1e-5}


@pytest.mark.parametrize("prec", ["c64", "c32"])
@pytest.mark.parametrize("rates", RATES)
def test_noisy_unitary_exact_bit_identical(prec, rates):
    n = 9
    flush("noisy unitary exact %s %s" % (prec, rates))
    c = mixed_circuit(n, 4, 21)
    draws = workloads.uniforms(99, 8 * len(c.gates()) + 8)
    res = Simulator.run_with_config(c, SimConfig(precision=prec, noise=NoiseModel(*rates), draws=draws))
    want, _, cnt, used = orc.run_circuit_noisy(n, c.to_array(), *rates, draws, f32=(prec == "c32"))
    assert res.draws_used == used
    assert np.array_equal(res.state.state_vector(), want)
    assert res.metrics.gate_count == cnt


@pytest.mark.parametrize("prec", ["c64", "c32"])
@pytest.mark.parametrize("rates", RATES)
def test_noisy_measured_exact(prec, rates):
    n = 9
    flush("noisy measured exact %s %s" % (prec, rates))
    c = noisy_circuit(n, 21)
    draws = workloads.uniforms(99, 8 * len(c.gates()) + 8)
    res = Simulator.run_with_config(c, SimConfig(precision=prec, noise=NoiseModel(*rates), draws=draws))
    want, wouts, cnt, used = orc.run_circuit_noisy(n, c.to_array(), *rates, draws, f32=(prec == "c32"))
    assert res.draws_used == used
    assert max_abs(res.state.state_vector(), want) < TOL[prec]
    assert [(m.qubit, int(m.result)) for m in res.measurements] == [(int(o[0]), int(o[1])) for o in wouts]
    assert all(abs(m.probability - o[2]) < TOL[prec] for m, o in zip(res.measurements, wouts))
    assert res.metrics.gate_count == cnt


@pytest.mark.parametrize("prec", ["c64", "c32"])
def test_noisy_fused_within_tolerance(prec):
    n = 12
    flush("noisy fused %s" % prec)
    c = noisy_circuit(n, 5)
    rates = (0.05, 0.1, 0.1)
    draws = workloads.uniforms(7, 8 * len(c.gates()) + 8)
    res = Simulator.run_with_config(c, SimConfig(precision=prec, fusion=True, noise=NoiseModel(*rates), draws=draws))
    want, _, cnt, used = orc.run_circuit_noisy(n, c.to_array(), *rates, draws)
    assert res.draws_used == used
    assert max_abs(res.state.state_vector(), want) < (1e-12 if prec == "c64" else 1e-5)
    assert res.metrics.gate_count == cnt


def test_noisy_run_draws_its_own_uniforms():
    c = noisy_circuit(6, 2)
    nm = NoiseModel(0.1, 0.1, 0.1)
    res = Simulator.run_with_config(c, SimConfig(seed=3, noise=nm))
    assert 0 < res.draws_used <= Simulator.noise_draw_bound(c, nm)
    assert abs(res.state.norm_sq() - 1.0) < 1e-12


@pytest.mark.parametrize("prec", ["c64", "c32"])
def test_checkpoint_bytes_and_restore(prec):
    n = 11
    flush("checkpoint %s" % prec)
    c = mixed_circuit(n, 3, 8)
    st = Simulator.run_with_config(c, SimConfig(precision=prec)).state
    sv = st.state_vector()
    cp = st.checkpoint()
    assert cp.size_bytes() == 16 << n and cp.num_amplitudes == 1 << n
    assert cp.data == StateCheckpoint.capture(sv.astype(np.complex128)).data
    back = QuantumState.from_checkpoint(cp, prec)
    assert np.array_equal(back.state_vector(), sv)
    assert np.array_equal(cp.restore(), sv.astype(np.complex128))


def test_checkpoint_across_precisions_and_sizes():
    n = 24  # 2^24 amplitudes: several staging chunks
    flush("checkpoint 24q")
    c = workloads.random_circuit(n, 2, 4)
    st = Simulator.run_with_config(c, SimConfig(precision="c64", fusion=True)).state
    cp = st.checkpoint()
    sv = st.state_vector()
    f32 = QuantumState.from_checkpoint(cp, "c32")
    assert np.array_equal(f32.state_vector(), sv.astype(np.complex64))
    cp32 = f32.checkpoint()
    assert np.array_equal(np.frombuffer(cp32.data, dtype="<c16"), sv.astype(np.complex64).astype(np.complex128))
    f64 = QuantumState.from_checkpoint(cp.data, "c64")
    assert np.array_equal(f64.state_vector(), sv)


def test_checkpoint_rejects_bad_sizes():
    for nbytes in (15, 48, 0):
        with pytest.raises(InvalidStateVector):
            QuantumState.from_checkpoint(bytes(nbytes), "c64")
    one = QuantumState.from_checkpoint(StateCheckpoint.capture(np.array([1.0 + 0j])), "c64")  # from_amplitudes(_, 0)
    assert one.num_qubits() == 0
