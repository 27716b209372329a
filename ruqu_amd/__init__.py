"""ruqu_amd: MI355X-native state-vector engine with ruqu-core's circuit/state API.

The compute path is libruqu_gpu.so (hand-written HIP kernels for gfx950 and a
C++ runtime behind a C ABI, include/ruqu_gpu.h).  This package is its Python
binding, mirroring the Rust names (QuantumState, Simulator, QuantumCircuit,
Gate, PauliString, Hamiltonian) so host code and tests read like the
reference's own.
"""
from .gates import Gate, GateKind, Hamiltonian, PauliOp, PauliString, QuantumCircuit
from .rng import StdRng
from .state import (
    ArgumentError,
    CircuitError,
    DeviceError,
    InvalidQubitIndex,
    InvalidStateVector,
    MeasurementOutcome,
    MemoryAllocationFailed,
    NoiseModel,
    QuantumError,
    QuantumState,
    QuantumStateF32,
    QubitLimitExceeded,
    ShotResult,
    SimConfig,
    SimulationMetrics,
    SimulationResult,
    Simulator,
    StateCheckpoint,
)

__all__ = [
    "Gate", "GateKind", "Hamiltonian", "PauliOp", "PauliString", "QuantumCircuit", "ArgumentError", "CircuitError",
    "DeviceError", "InvalidQubitIndex", "InvalidStateVector", "MeasurementOutcome", "MemoryAllocationFailed",
    "NoiseModel", "QuantumError", "QuantumState", "QuantumStateF32", "QubitLimitExceeded", "ShotResult", "SimConfig",
    "SimulationMetrics", "SimulationResult", "Simulator", "StateCheckpoint", "StdRng",
]
