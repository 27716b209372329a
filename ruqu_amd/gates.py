"""Gate IR and circuit builder: gate.rs, circuit.rs and the Pauli containers of
types.rs, as plain host data.  A circuit is an ordered list of Gate values;
`QuantumCircuit.to_array()` packs it into the rq_gate records the C ABI takes.
"""
import enum

import numpy as np

from ._native import GATE_DTYPE, PAULI_DTYPE


class GateKind(enum.IntEnum):
    """gate.rs:8-36 in declaration order (the rq_gate_kind values)."""

    H = 0
    X = 1
    Y = 2
    Z = 3
    S = 4
    Sdg = 5
    T = 6
    Tdg = 7
    Rx = 8
    Ry = 9
    Rz = 10
    Phase = 11
    CNOT = 12
    CZ = 13
    SWAP = 14
    Rzz = 15
    Measure = 16
    Reset = 17
    Barrier = 18
    Unitary1Q = 19


_TWO_Q = {GateKind.CNOT, GateKind.CZ, GateKind.SWAP, GateKind.Rzz}
_NON_UNITARY = {GateKind.Measure, GateKind.Reset, GateKind.Barrier}


class Gate:
    """One `Gate` value (gate.rs:8-36).  Build with the class methods named like
    the Rust variants: Gate.H(q), Gate.Rx(q, theta), Gate.CNOT(c, t),
    Gate.Unitary1Q(q, m), Gate.Barrier()."""

    __slots__ = ("kind", "q0", "q1", "theta", "u")

    def __init__(self, kind, q0=0, q1=0, theta=0.0, u=None):
        self.kind = GateKind(kind)
        self.q0 = int(q0)
        self.q1 = int(q1)
        self.theta = float(theta)
        self.u = None if u is None else np.asarray(u, dtype=np.complex128).reshape(2, 2)

    def qubits(self):  # gate.rs:40-64
        if self.kind == GateKind.Barrier:
            return []
        if self.kind in _TWO_Q:
            return [self.q0, self.q1]
        return [self.q0]

    def is_non_unitary(self):  # gate.rs:67-69
        return self.kind in _NON_UNITARY

    def __repr__(self):
        k = self.kind.name
        if self.kind == GateKind.Barrier:
            return "Barrier"
        if self.kind in _TWO_Q:
            extra = ", %r" % self.theta if self.kind == GateKind.Rzz else ""
            return "%s(%d, %d%s)" % (k, self.q0, self.q1, extra)
        if self.kind in (GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.Phase):
            return "%s(%d, %r)" % (k, self.q0, self.theta)
        return "%s(%d)" % (k, self.q0)


def _mk1(kind):
    return classmethod(lambda cls, q: cls(kind, q))


def _mk1a(kind):
    return classmethod(lambda cls, q, theta: cls(kind, q, 0, theta))


def _mk2(kind):
    return classmethod(lambda cls, a, b: cls(kind, a, b))


for _k in (GateKind.H, GateKind.X, GateKind.Y, GateKind.Z, GateKind.S, GateKind.Sdg, GateKind.T, GateKind.Tdg,
           GateKind.Measure, GateKind.Reset):
    setattr(Gate, _k.name, _mk1(_k))
for _k in (GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.Phase):
    setattr(Gate, _k.name, _mk1a(_k))
for _k in (GateKind.CNOT, GateKind.CZ, GateKind.SWAP):
    setattr(Gate, _k.name, _mk2(_k))
Gate.Rzz = classmethod(lambda cls, a, b, theta: cls(GateKind.Rzz, a, b, theta))
Gate.Barrier = classmethod(lambda cls: cls(GateKind.Barrier))
Gate.Unitary1Q = classmethod(lambda cls, q, m: cls(GateKind.Unitary1Q, q, 0, 0.0, m))


def gates_to_array(gates):
    """Pack Gate values into rq_gate records."""
    arr = np.zeros(len(gates), dtype=GATE_DTYPE)
    for i, g in enumerate(gates):
        arr[i]["kind"] = int(g.kind)
        arr[i]["q0"] = g.q0
        arr[i]["q1"] = g.q1
        arr[i]["theta"] = g.theta
        if g.u is not None:
            arr[i]["u"] = np.stack([g.u.real, g.u.imag], axis=-1).reshape(8)
    return arr


class QuantumCircuit:
    """circuit.rs:8-185: ordered gates on `num_qubits` qubits, fluent builder."""

    def __init__(self, num_qubits):
        self._gates = []
        self._num_qubits = int(num_qubits)
        self._array = None

    def _push(self, g):
        self._gates.append(g)
        self._array = None
        return self

    def h(self, q): return self._push(Gate.H(q))
    def x(self, q): return self._push(Gate.X(q))
    def y(self, q): return self._push(Gate.Y(q))
    def z(self, q): return self._push(Gate.Z(q))
    def s(self, q): return self._push(Gate.S(q))
    def t(self, q): return self._push(Gate.T(q))
    def rx(self, q, a): return self._push(Gate.Rx(q, a))
    def ry(self, q, a): return self._push(Gate.Ry(q, a))
    def rz(self, q, a): return self._push(Gate.Rz(q, a))
    def phase(self, q, a): return self._push(Gate.Phase(q, a))
    def cnot(self, c, t): return self._push(Gate.CNOT(c, t))
    def cz(self, a, b): return self._push(Gate.CZ(a, b))
    def swap(self, a, b): return self._push(Gate.SWAP(a, b))
    def rzz(self, a, b, ang): return self._push(Gate.Rzz(a, b, ang))
    def measure(self, q): return self._push(Gate.Measure(q))
    def reset(self, q): return self._push(Gate.Reset(q))
    def barrier(self): return self._push(Gate.Barrier())
    def add_gate(self, g): return self._push(g)

    def measure_all(self):  # circuit.rs:110-115
        for q in range(self._num_qubits):
            self._push(Gate.Measure(q))
        return self

    def gates(self):
        return list(self._gates)

    def num_qubits(self):
        return self._num_qubits

    def gate_count(self):
        return len(self._gates)

    def to_array(self):
        if self._array is None:
            self._array = gates_to_array(self._gates)
        return self._array

    @classmethod
    def from_array(cls, num_qubits, arr):
        c = cls(num_qubits)
        for r in arr:
            u = None
            if int(r["kind"]) == GateKind.Unitary1Q:
                u = (r["u"][0::2] + 1j * r["u"][1::2]).reshape(2, 2)
            c._gates.append(Gate(int(r["kind"]), int(r["q0"]), int(r["q1"]), float(r["theta"]), u))
        c._array = np.array(arr, dtype=GATE_DTYPE)
        return c

    def depth(self):  # circuit.rs:153-184
        from ._native import lib, ptr
        arr = self.to_array()
        return int(lib().rq_circuit_depth(self._num_qubits, ptr(arr), len(arr)))


class PauliOp(enum.IntEnum):
    """types.rs:190-195."""

    I = 0  # noqa: E741
    X = 1
    Y = 2
    Z = 3


class PauliString:
    """types.rs:201-213: sparse (qubit, PauliOp) factors."""

    def __init__(self, ops):
        self.ops = [(int(q), PauliOp(op)) for q, op in ops]

    @classmethod
    def identity(cls):
        return cls([])

    def to_array(self):
        a = np.zeros(len(self.ops), dtype=PAULI_DTYPE)
        for i, (q, op) in enumerate(self.ops):
            a[i] = (q, int(op))
        return a


class Hamiltonian:
    """types.rs:216-226: sum of c_k * P_k."""

    def __init__(self, terms, num_qubits):
        self.terms = [(float(c), p if isinstance(p, PauliString) else PauliString(p)) for c, p in terms]
        self.num_qubits = int(num_qubits)

    def to_arrays(self):
        coeffs = np.array([c for c, _ in self.terms], dtype=np.float64)
        offs = np.zeros(len(self.terms) + 1, dtype=np.uint64)
        parts = []
        for k, (_, p) in enumerate(self.terms):
            a = p.to_array()
            parts.append(a)
            offs[k + 1] = offs[k] + len(a)
        ops = np.concatenate(parts) if parts else np.zeros(0, dtype=PAULI_DTYPE)
        return coeffs, offs, ops
