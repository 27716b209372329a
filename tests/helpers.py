"""Shared test helpers: dense reference matrices built from first principles
(numpy Kronecker products), random states, and comparison utilities."""
import numpy as np

from oracle import oracle as orc
from ruqu_amd.gates import GateKind, QuantumCircuit

I2 = np.eye(2, dtype=np.complex128)
PAULI = {
    0: I2,
    1: np.array([[0, 1], [1, 0]], dtype=np.complex128),
    2: np.array([[0, -1j], [1j, 0]], dtype=np.complex128),
    3: np.array([[1, 0], [0, -1]], dtype=np.complex128),
}


def embed_1q(n, q, m):
    """Full 2^n x 2^n operator of a 2x2 on qubit q (qubit q = bit q)."""
    full = np.array([[1.0 + 0j]])
    for k in range(n - 1, -1, -1):
        full = np.kron(full, m if k == q else I2)
    return full


def embed_2q(n, q1, q2, m4):
    """Full operator of a 4x4 whose row/col index is q1_bit*2 + q2_bit."""
    dim = 1 << n
    out = np.zeros((dim, dim), dtype=np.complex128)
    for col in range(dim):
        b1, b2 = (col >> q1) & 1, (col >> q2) & 1
        c = b1 * 2 + b2
        for r in range(4):
            row = col & ~((1 << q1) | (1 << q2))
            row |= ((r >> 1) & 1) << q1
            row |= (r & 1) << q2
            out[row, col] += m4[r, c]
    return out


def pauli_matrix(n, ops):
    full = np.eye(1 << n, dtype=np.complex128)
    for q, op in ops:
        full = embed_1q(n, q, PAULI[op]) @ full
    return full


def random_state(n, seed, dtype=np.complex128):
    rng = np.random.default_rng(seed)
    v = rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n)
    v /= np.linalg.norm(v)
    return v.astype(dtype)


def random_unitary(k, seed):
    rng = np.random.default_rng(seed)
    z = rng.normal(size=(k, k)) + 1j * rng.normal(size=(k, k))
    q, r = np.linalg.qr(z)
    return q * (np.diag(r) / np.abs(np.diag(r)))


ALL_1Q = [GateKind.H, GateKind.X, GateKind.Y, GateKind.Z, GateKind.S, GateKind.Sdg, GateKind.T, GateKind.Tdg,
          GateKind.Rx, GateKind.Ry, GateKind.Rz, GateKind.Phase]
ALL_2Q = [GateKind.CNOT, GateKind.CZ, GateKind.SWAP, GateKind.Rzz]


def mixed_circuit(n, layers, seed, with_unitary=True):
    """Every gate variant on random qubits: a broad parity workload."""
    rng = np.random.default_rng(seed)
    c = QuantumCircuit(n)
    from ruqu_amd.gates import Gate
    for _ in range(layers):
        for q in range(n):
            k = ALL_1Q[rng.integers(len(ALL_1Q))]
            c.add_gate(Gate(k, q, 0, rng.uniform(0, 2 * np.pi)))
        if with_unitary:
            c.add_gate(Gate.Unitary1Q(int(rng.integers(n)), random_unitary(2, int(rng.integers(1 << 30)))))
        perm = rng.permutation(n)
        for j in range(n // 2):
            k = ALL_2Q[rng.integers(len(ALL_2Q))]
            c.add_gate(Gate(k, int(perm[2 * j]), int(perm[2 * j + 1]), rng.uniform(0, 2 * np.pi)))
    return c


def max_abs(a, b):
    return float(np.max(np.abs(np.asarray(a, dtype=np.complex128) - np.asarray(b, dtype=np.complex128))))


__all__ = ["orc", "embed_1q", "embed_2q", "pauli_matrix", "random_state", "random_unitary", "mixed_circuit",
           "max_abs", "ALL_1Q", "ALL_2Q", "PAULI"]
