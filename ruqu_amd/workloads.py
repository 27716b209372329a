"""Seeded synthetic workloads (the reference ships no generator).

Stream: SplitMix64(seed) -> u64; uniform = (u64 >> 11) * 2^-53.

random_circuit(n, depth, seed): every layer applies Rx, Ry and Rz (in that
order) to every qubit with angles 2*pi*U(0,1), then a Fisher-Yates shuffle of
the qubits picks min(10, n//2) disjoint pairs (perm[0],perm[1]), (perm[2],
perm[3]), ...; each pair gets CNOT(first, second) when the next uniform is
< 0.5, else CZ(first, second).

Graphs: random d-regular by the configuration model with rejection of
self-loops / multi-edges; G(n, p) by one uniform per pair (i<j) in
lexicographic order; weights U(0,1).
"""
import math

import numpy as np

from .gates import Gate, GateKind, QuantumCircuit

MASK64 = (1 << 64) - 1


class SplitMix64:
    def __init__(self, seed):
        self.state = seed & MASK64

    def next_u64(self):
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / 9007199254740992.0)

    def below(self, k):
        """Integer in [0, k) from one uniform."""
        return min(int(self.uniform() * k), k - 1)

    def normal(self):
        """Box-Muller from two uniforms."""
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def uniforms(seed, count):
    """`count` uniforms of SplitMix64(seed), vectorised."""
    idx = np.arange(1, count + 1, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = np.uint64(seed & MASK64) + idx * np.uint64(0x9E3779B97F4A7C15)
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def stream_seed(base, index):
    """Seed of instance / shot `index` of a workload with base seed `base`."""
    return ((base & 0xFFFFFFFF) << 32) | (index & 0xFFFFFFFF)


def shot_seed(base):
    return ((base & 0xFFFFFFFF) << 32) | 0x80000000


def random_circuit(n, depth, seed, pairs=10):
    rng = SplitMix64(seed)
    c = QuantumCircuit(n)
    two_pi = 2.0 * math.pi
    npairs = min(pairs, n // 2)
    for _ in range(depth):
        for q in range(n):
            c.rx(q, two_pi * rng.uniform())
            c.ry(q, two_pi * rng.uniform())
            c.rz(q, two_pi * rng.uniform())
        perm = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            perm[i], perm[j] = perm[j], perm[i]
        for k in range(npairs):
            a, b = perm[2 * k], perm[2 * k + 1]
            if rng.uniform() < 0.5:
                c.cnot(a, b)
            else:
                c.cz(a, b)
    return c


def random_regular_graph(n, d, seed, max_tries=10000):
    """Edges (i<j) of a random d-regular graph (configuration model)."""
    rng = SplitMix64(seed)
    for _ in range(max_tries):
        stubs = [v for v in range(n) for _ in range(d)]
        for i in range(len(stubs) - 1, 0, -1):
            j = rng.below(i + 1)
            stubs[i], stubs[j] = stubs[j], stubs[i]
        edges = set()
        ok = True
        for k in range(0, len(stubs), 2):
            a, b = stubs[k], stubs[k + 1]
            if a == b or (min(a, b), max(a, b)) in edges:
                ok = False
                break
            edges.add((min(a, b), max(a, b)))
        if ok:
            return sorted(edges)
    raise RuntimeError("no simple %d-regular graph found" % d)


def erdos_renyi_graph(n, p, seed, weighted=True):
    """(edges, weights) of G(n, p); weights U(0,1) when weighted else 1."""
    rng = SplitMix64(seed)
    edges, w = [], []
    for i in range(n):
        for j in range(i + 1, n):
            if rng.uniform() < p:
                edges.append((i, j))
                w.append(rng.uniform() if weighted else 1.0)
    return edges, w


def random_pauli_hamiltonian(n, terms, max_weight, seed):
    """`terms` Pauli strings of weight 1..max_weight on distinct qubits,
    coefficients N(0,1).  Returns [(coeff, [(qubit, op), ...]), ...]."""
    rng = SplitMix64(seed)
    out = []
    for _ in range(terms):
        w = 1 + rng.below(max_weight)
        qs = list(range(n))
        for i in range(n - 1, 0, -1):
            j = rng.below(i + 1)
            qs[i], qs[j] = qs[j], qs[i]
        ops = sorted((qs[k], 1 + rng.below(3)) for k in range(w))
        out.append((rng.normal(), ops))
    return out


def ry_cz_ansatz(n, layers, params):
    """Hardware-efficient Ry-CZ ansatz: per layer Ry(theta) on every qubit, then
    CZ on the chain (0,1), (1,2), ...; len(params) = n * layers."""
    c = QuantumCircuit(n)
    k = 0
    for _ in range(layers):
        for q in range(n):
            c.ry(q, params[k])
            k += 1
        for q in range(n - 1):
            c.cz(q, q + 1)
    return c


__all__ = ["SplitMix64", "uniforms", "stream_seed", "shot_seed", "random_circuit", "random_regular_graph",
           "erdos_renyi_graph", "random_pauli_hamiltonian", "ry_cz_ansatz", "Gate", "GateKind"]
