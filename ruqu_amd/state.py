"""QuantumState / QuantumStateF32 / Simulator over the C ABI.

Mirrors the Rust API of state.rs, mixed_precision.rs and simulator.rs (same
method names, argument meaning and error variants).  Amplitudes live in GPU
memory owned by the native handle; state_vector() and set_amplitudes() are
explicit copies.

Randomness: the reference draws measurement uniforms from a seeded StdRng
inside the state (state.rs:302).  Here every measuring call accepts explicit
draws; when none are given they come from the same stream: rand 0.8.6's
StdRng seeded with the state's seed (ruqu_amd.rng, restated in the native
library), or OS entropy for an unseeded state.  Seeded runs therefore
reproduce the reference's outcomes.
"""
import ctypes as C
import time

import numpy as np

from . import _native
from ._native import MEAS_DTYPE, Measurement, Metrics, Noise, lib, ptr
from .rng import shot_draws, state_rng
from .gates import GateKind, QuantumCircuit

C64, C32 = 0, 1
_PREC = {"c64": C64, "f64": C64, C64: C64, "c32": C32, "f32": C32, C32: C32}


class QuantumError(Exception):
    """error.rs:8-23."""

    code = None


class QubitLimitExceeded(QuantumError):
    code = 1


class InvalidQubitIndex(QuantumError):
    code = 2


class MemoryAllocationFailed(QuantumError):
    code = 3


class InvalidStateVector(QuantumError):
    code = 4


class CircuitError(QuantumError):
    code = 5


class DeviceError(QuantumError):
    code = 6


class ArgumentError(QuantumError):
    code = 7


_ERRORS = {c.code: c for c in (QubitLimitExceeded, InvalidQubitIndex, MemoryAllocationFailed, InvalidStateVector,
                               CircuitError, DeviceError, ArgumentError)}


def check(rc):
    if rc != 0:
        msg = lib().rq_last_error()
        raise _ERRORS.get(rc, QuantumError)(msg.decode() if msg else "error %d" % rc)


class MeasurementOutcome:
    """types.rs:230-234."""

    __slots__ = ("qubit", "result", "probability")

    def __init__(self, qubit, result, probability):
        self.qubit = int(qubit)
        self.result = bool(result)
        self.probability = float(probability)

    def __repr__(self):
        return "MeasurementOutcome(qubit=%d, result=%s, probability=%r)" % (self.qubit, self.result, self.probability)


class SimulationMetrics:
    """types.rs:238-245 plus device figures (passes, device_time_ms, hbm_bytes)."""

    def __init__(self, m):
        for name, _ in Metrics._fields_:
            if not name.startswith("_"):
                setattr(self, name, getattr(m, name))


def _as_prec(precision):
    try:
        return _PREC[precision]
    except KeyError:
        raise ValueError("precision must be 'c64' or 'c32'")


class QuantumState:
    """state.rs QuantumState (precision 'c64') or mixed_precision.rs
    QuantumStateF32 (precision 'c32'), resident on the GPU."""

    MAX_QUBITS = 32  # state.rs:15 (the device limit is rq_max_qubits)

    def __init__(self, handle, precision, seed=None):
        self._h = C.c_void_p(handle)
        self._prec = precision
        self._rng = state_rng(seed)

    # ---- construction (state.rs:31-96) --------------------------------------
    @classmethod
    def new(cls, num_qubits, precision="c64"):
        p = _as_prec(precision)
        h = C.c_void_p()
        check(lib().rq_state_create(int(num_qubits), p, 0, C.byref(h)))
        return cls(h.value, p)

    @classmethod
    def new_with_seed(cls, num_qubits, seed, precision="c64"):
        p = _as_prec(precision)
        h = C.c_void_p()
        check(lib().rq_state_create(int(num_qubits), p, 1, C.byref(h)))
        return cls(h.value, p, seed)

    @classmethod
    def from_amplitudes(cls, amps, num_qubits, precision=None):
        amps = np.asarray(amps)
        if precision is None:
            precision = C32 if amps.dtype == np.complex64 else C64
        p = _as_prec(precision)
        a = np.ascontiguousarray(amps, dtype=np.complex128 if p == C64 else np.complex64)
        h = C.c_void_p()
        check(lib().rq_state_from_amplitudes(ptr(a), a.size, int(num_qubits), p, C.byref(h)))
        return cls(h.value, p)

    @classmethod
    def from_checkpoint(cls, checkpoint, precision="c64"):
        """A new device state from StateCheckpoint bytes (replay.rs:270-287,
        then from_amplitudes); 'c32' narrows each f64 to nearest."""
        data = checkpoint.data if isinstance(checkpoint, StateCheckpoint) else bytes(checkpoint)
        p = _as_prec(precision)
        buf = np.frombuffer(data, dtype=np.uint8)
        h = C.c_void_p()
        check(lib().rq_state_from_checkpoint(ptr(buf) if buf.size else None, buf.size, p, C.byref(h)))
        return cls(h.value, p)

    def checkpoint(self):
        """StateCheckpoint::capture of this state's amplitudes (replay.rs:257-268):
        little-endian f64 (re, im) pairs, complex-f32 widened exactly."""
        out = np.empty(int(lib().rq_state_checkpoint_bytes(self._h)), dtype=np.uint8)
        check(lib().rq_state_checkpoint(self._h, ptr(out)))
        return StateCheckpoint(out.tobytes(), out.size // 16)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rq_state_destroy(h)
            except Exception:
                pass
            self._h = None

    def close(self):
        self.__del__()

    @property
    def handle(self):
        return self._h

    @property
    def precision(self):
        return "c64" if self._prec == C64 else "c32"

    # ---- accessors ----------------------------------------------------------
    def num_qubits(self):
        return int(lib().rq_state_num_qubits(self._h))

    def num_amplitudes(self):
        return 1 << self.num_qubits()

    def state_vector(self):
        out = np.empty(self.num_amplitudes(), dtype=np.complex128 if self._prec == C64 else np.complex64)
        check(lib().rq_state_read(self._h, ptr(out)))
        return out

    def set_amplitudes(self, amps):
        """Write-back counterpart of amplitudes_mut() (state.rs:118-120)."""
        a = np.ascontiguousarray(amps, dtype=np.complex128 if self._prec == C64 else np.complex64)
        if a.size != self.num_amplitudes():
            raise InvalidStateVector("invalid state vector: length %d does not match 2^%d" % (a.size, self.num_qubits()))
        check(lib().rq_state_write(self._h, ptr(a)))

    def probabilities(self):
        out = np.empty(self.num_amplitudes(), dtype=np.float64)
        check(lib().rq_probabilities(self._h, ptr(out)))
        return out

    def probability_of_qubit(self, qubit):
        v = C.c_double()
        check(lib().rq_probability_of_qubit(self._h, int(qubit), C.byref(v)))
        return v.value

    def measurement_record(self):
        n = lib().rq_measurement_record(self._h, None, 0)
        rec = np.zeros(n, dtype=MEAS_DTYPE)
        lib().rq_measurement_record(self._h, ptr(rec), n)
        return [MeasurementOutcome(r["qubit"], r["result"], r["probability"]) for r in rec]

    @staticmethod
    def estimate_memory(num_qubits, precision="c64"):
        return int(lib().rq_estimate_memory(int(num_qubits), _as_prec(precision)))

    def precision_error_bound(self):
        return float(lib().rq_precision_error_bound(self._h))

    @classmethod
    def from_f64(cls, state):
        """QuantumStateF32::from_f64 (mixed_precision.rs:260-278): a new
        complex-f32 state, each amplitude of the complex-f64 `state` rounded to
        nearest, with its measurement record."""
        h = C.c_void_p()
        check(lib().rq_state_from_f64(state.handle, C.byref(h)))
        return cls(h.value, C32)

    def to_f64(self):
        """QuantumStateF32::to_f64 (mixed_precision.rs:280-288): a new
        complex-f64 state, each amplitude widened exactly, no measurement
        record."""
        h = C.c_void_p()
        check(lib().rq_state_to_f64(self._h, C.byref(h)))
        return QuantumState(h.value, C64)

    def copy(self):
        h = C.c_void_p()
        check(lib().rq_state_create(self.num_qubits(), self._prec, 1, C.byref(h)))
        other = QuantumState(h.value, self._prec)
        check(lib().rq_state_copy(self._h, other._h))
        return other

    # ---- gates --------------------------------------------------------------
    def _draw(self, draw):
        return float(self._rng.random()) if draw is None else float(draw)

    def apply_gate(self, gate, draw=None):
        """state.rs:158-204; returns the list of measurement outcomes."""
        arr = QuantumCircuit(0).add_gate(gate).to_array()
        needs = gate.kind in (GateKind.Measure, GateKind.Reset)
        m = Measurement()
        has = C.c_int(0)
        check(lib().rq_apply_gate(self._h, ptr(arr), self._draw(draw) if needs else 0.0, C.byref(m), C.byref(has)))
        return [MeasurementOutcome(m.qubit, m.result, m.probability)] if has.value else []

    def apply_single_qubit_gate(self, qubit, matrix):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(2, 2))
        check(lib().rq_apply_1q(self._h, int(qubit), ptr(m)))

    def apply_two_qubit_gate(self, q1, q2, matrix):
        m = np.ascontiguousarray(np.asarray(matrix, dtype=np.complex128).reshape(4, 4))
        check(lib().rq_apply_2q(self._h, int(q1), int(q2), ptr(m)))

    # ---- measurement --------------------------------------------------------
    def measure(self, qubit, draw=None):
        m = Measurement()
        check(lib().rq_measure(self._h, int(qubit), self._draw(draw), C.byref(m)))
        return MeasurementOutcome(m.qubit, m.result, m.probability)

    def measure_all(self, draws=None):
        n = self.num_qubits()
        d = self._rng.random(n) if draws is None else np.ascontiguousarray(draws, dtype=np.float64)
        out = np.zeros(n, dtype=MEAS_DTYPE)
        check(lib().rq_measure_all(self._h, ptr(d), ptr(out)))
        return [MeasurementOutcome(r["qubit"], r["result"], r["probability"]) for r in out]

    def reset_qubit(self, qubit, draw=None):
        check(lib().rq_reset_qubit(self._h, int(qubit), self._draw(draw)))

    def sample_shots(self, draws):
        """measure_all on copies of the state, one row of draws per shot.
        Returns (bits[shots, n] as uint8, near_ties)."""
        n = self.num_qubits()
        d = np.ascontiguousarray(draws, dtype=np.float64).reshape(-1, n)
        bits = np.zeros(d.shape, dtype=np.uint8)
        ties = C.c_uint64(0)
        check(lib().rq_sample_shots(self._h, ptr(d), d.shape[0], ptr(bits), C.byref(ties)))
        return bits, int(ties.value)

    # ---- observables ----------------------------------------------------------
    def expectation_value(self, pauli):
        a = pauli.to_array()
        v = C.c_double()
        check(lib().rq_expectation(self._h, ptr(a), len(a), C.byref(v)))
        return v.value

    def expectation_hamiltonian(self, h):
        coeffs, offs, ops = h.to_arrays()
        v = C.c_double()
        check(lib().rq_expectation_hamiltonian(self._h, ptr(coeffs), ptr(offs), ptr(ops), len(coeffs), C.byref(v)))
        return v.value

    def expectation_z_all(self):
        out = np.zeros(self.num_qubits(), dtype=np.float64)
        check(lib().rq_expectation_z_all(self._h, ptr(out)))
        return out

    def set_zero(self):
        """Back to |0...0> in place (the state QuantumState::new starts in)."""
        check(lib().rq_state_set_zero(self._h))

    def normalize(self):
        check(lib().rq_normalize(self._h))

    def norm_sq(self):
        v = C.c_double()
        check(lib().rq_norm_sq(self._h, C.byref(v)))
        return v.value

    def fidelity(self, other):
        v = C.c_double()
        check(lib().rq_fidelity(self._h, other._h, C.byref(v)))
        return v.value


QuantumStateF32 = QuantumState  # precision='c32' selects the f32 twin


class StateCheckpoint:
    """replay.rs:251-294: a serialised snapshot of a state vector, amplitudes
    as interleaved little-endian f64 (re, im) pairs, 16 bytes each."""

    def __init__(self, data, num_amplitudes):
        self.data = bytes(data)
        self.num_amplitudes = int(num_amplitudes)

    @classmethod
    def capture(cls, amplitudes):  # replay.rs:257-268
        a = np.ascontiguousarray(amplitudes, dtype=np.complex128)
        return cls(a.astype("<c16").tobytes(), a.size)

    def restore(self):  # replay.rs:270-287
        if len(self.data) < 16 * self.num_amplitudes:
            raise ValueError("checkpoint data corrupted")
        return np.frombuffer(self.data, dtype="<c16", count=self.num_amplitudes).astype(np.complex128)

    def size_bytes(self):  # replay.rs:290-293
        return len(self.data)


class NoiseModel:
    """types.rs:249-253."""

    def __init__(self, depolarizing_rate=0.0, bit_flip_rate=0.0, phase_flip_rate=0.0):
        self.depolarizing_rate = depolarizing_rate
        self.bit_flip_rate = bit_flip_rate
        self.phase_flip_rate = phase_flip_rate


class SimConfig:
    """simulator.rs:14-31 (+ precision, fusion and explicit draws)."""

    def __init__(self, seed=None, noise=None, shots=None, precision="c64", fusion=False, draws=None):
        self.seed = seed
        self.noise = noise
        self.shots = shots
        self.precision = precision
        self.fusion = fusion
        self.draws = draws


class SimulationResult:
    def __init__(self, state, measurements, metrics):
        self.state = state
        self.measurements = measurements
        self.metrics = metrics


class ShotResult:
    def __init__(self, counts, metrics):
        self.counts = counts
        self.metrics = metrics


def _count_draws(arr):
    k = arr["kind"]
    return int(np.count_nonzero((k == GateKind.Measure) | (k == GateKind.Reset)))


def _fusion_mode(f):
    """False/0: exact per-gate kernels; True/1/"fused": the fused tile
    schedule; 2/"dense": dense 5-qubit blocks on the matrix cores."""
    if isinstance(f, str):
        return {"exact": 0, "fused": 1, "dense": 2}[f]
    return int(f)


class Simulator:
    """simulator.rs:47-164 (stateless)."""

    @staticmethod
    def run(circuit):
        return Simulator.run_with_config(circuit, SimConfig())

    @staticmethod
    def run_with_config(circuit, config, state=None):
        noise = config.noise
        if noise is not None and any(
                getattr(noise, f) > 0 for f in ("depolarizing_rate", "bit_flip_rate", "phase_flip_rate")):
            return Simulator._run_noisy(circuit, config, state)
        arr = circuit.to_array()
        p = _as_prec(config.precision)
        if state is None:
            if config.seed is not None:
                state = QuantumState.new_with_seed(circuit.num_qubits(), config.seed, p)
            else:
                state = QuantumState.new(circuit.num_qubits(), p)
        nd = _count_draws(arr)
        if config.draws is not None:
            draws = np.ascontiguousarray(config.draws, dtype=np.float64)
        else:
            draws = state._rng.random(nd)
        meas = np.zeros(max(nd, 1), dtype=MEAS_DTYPE)
        nout = C.c_size_t(0)
        met = Metrics()
        check(lib().rq_run_circuit(state.handle, ptr(arr), len(arr), ptr(draws), len(draws), ptr(meas), len(meas),
                                   C.byref(nout), _fusion_mode(config.fusion), C.byref(met)))
        outs = [MeasurementOutcome(r["qubit"], r["result"], r["probability"]) for r in meas[: nout.value]]
        return SimulationResult(state, outs, SimulationMetrics(met))

    @staticmethod
    def noise_draw_bound(circuit, noise):
        """Most uniform draws a run with `noise` can consume: one per
        Measure/Reset plus, per qubit a gate touches, two for depolarising
        (when its rate > 0), one for bit flip and one for phase flip."""
        arr = circuit.to_array()
        k = arr["kind"]
        two = (k >= GateKind.CNOT) & (k <= GateKind.Rzz)
        touched = int(np.count_nonzero(k != GateKind.Barrier)) + int(np.count_nonzero(two))
        per = (2 if noise.depolarizing_rate > 0 else 0) + (noise.bit_flip_rate > 0) + (noise.phase_flip_rate > 0)
        return _count_draws(arr) + touched * per

    @staticmethod
    def _run_noisy(circuit, config, state=None):
        """simulator.rs:56-101 with config.noise (apply_noise, :178-221):
        one trajectory; config.draws (if given) is the stream of uniforms
        the reference's StdRng would supply, measurement and noise draws
        interleaved in the reference's order."""
        arr = circuit.to_array()
        p = _as_prec(config.precision)
        if state is None:
            if config.seed is not None:
                state = QuantumState.new_with_seed(circuit.num_qubits(), config.seed, p)
            else:
                state = QuantumState.new(circuit.num_qubits(), p)
        nz = config.noise
        if config.draws is not None:
            draws = np.ascontiguousarray(config.draws, dtype=np.float64)
        else:
            draws = state._rng.random(Simulator.noise_draw_bound(circuit, nz))
        nm = _count_draws(arr)
        meas = np.zeros(max(nm, 1), dtype=MEAS_DTYPE)
        nout, used = C.c_size_t(0), C.c_size_t(0)
        met = Metrics()
        cn = Noise(float(nz.depolarizing_rate), float(nz.bit_flip_rate), float(nz.phase_flip_rate))
        check(lib().rq_run_circuit_noisy(state.handle, ptr(arr), len(arr), C.byref(cn), ptr(draws), len(draws),
                                         C.byref(used), ptr(meas), len(meas), C.byref(nout),
                                         _fusion_mode(config.fusion), C.byref(met)))
        outs = [MeasurementOutcome(r["qubit"], r["result"], r["probability"]) for r in meas[: nout.value]]
        res = SimulationResult(state, outs, SimulationMetrics(met))
        res.draws_used = used.value
        return res

    @staticmethod
    def run_shots(circuit, shots, seed=None, precision="c64", fusion=False, draws=None):
        """simulator.rs:107-164.  Without Measure gates the circuit is simulated
        once and every shot is measure_all on a copy of the final state (the
        reference re-simulates per shot; the results are identical given the
        draws).  With Measure gates each shot is re-simulated as in the
        reference.  draws: optional [shots, k] uniforms per shot."""
        t0 = time.perf_counter()
        base = 42 if seed is None else seed  # simulator.rs:114
        n = circuit.num_qubits()
        arr = circuit.to_array()
        has_meas = bool(np.any(arr["kind"] == GateKind.Measure))
        nmr = _count_draws(arr)
        per_shot = nmr + (0 if has_meas else n)
        if draws is None:
            draws = shot_draws(base, shots, per_shot)  # shot s: StdRng(base + s), simulator.rs:122-127
        draws = np.ascontiguousarray(draws, dtype=np.float64).reshape(shots, per_shot)
        counts = {}
        gate_total = 0
        if not has_meas and nmr == 0:
            res = Simulator.run_with_config(circuit, SimConfig(seed=base, precision=precision, fusion=fusion))
            bits, _ = res.state.sample_shots(draws)
            gate_total = res.metrics.gate_count * shots
            keys, cnt = np.unique(bits, axis=0, return_counts=True)
            for k, c in zip(keys, cnt):
                counts[tuple(bool(b) for b in k)] = int(c)
        else:
            state = None
            for s in range(shots):
                cfg = SimConfig(seed=base + s, precision=precision, fusion=fusion, draws=draws[s, :nmr])
                res = Simulator.run_with_config(circuit, cfg, state=state)
                state = res.state
                gate_total += res.metrics.gate_count
                outs = list(res.measurements)
                if not has_meas:
                    outs += state.measure_all(draws[s, nmr:])
                bits = [False] * n
                for m in outs:
                    if m.qubit < n:
                        bits[m.qubit] = m.result
                key = tuple(bits)
                counts[key] = counts.get(key, 0) + 1
        el = time.perf_counter() - t0
        m = Metrics()
        m.num_qubits = n
        m.gate_count = gate_total
        m.execution_time_ns = int(el * 1e9)
        m.peak_memory_bytes = QuantumState.estimate_memory(n, precision)
        m.gates_per_second = gate_total / el if el > 0 else 0.0
        return ShotResult(counts, SimulationMetrics(m))
