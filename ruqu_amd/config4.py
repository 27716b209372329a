"""The batched workload: 4096 QAOA MaxCut instances (n ~ U{10..22}, G(n, 0.5),
weights U(0,1), p ~ U{1..5}) and 512 VQE instances (n ~ U{12..16}, Ry-CZ
ansatz with 4 layers, 100 Pauli strings of weight <= 4 with N(0,1)
coefficients), generated from SplitMix64 streams of base seed 4."""
from . import workloads
from .workloads import SplitMix64, stream_seed

BASE = 4


def qaoa_instances(count=4096, nmin=10, nmax=22, pmax=5, base=BASE):
    out = []
    for i in range(count):
        rng = SplitMix64(stream_seed(base, i))
        n = nmin + rng.below(nmax - nmin + 1)
        p = 1 + rng.below(pmax)
        edges, w = workloads.erdos_renyi_graph(n, 0.5, stream_seed(base, 1_000_000 + i))
        out.append((n, edges, w, p))
    return out


def vqe_instances(count=512, nmin=12, nmax=16, layers=4, terms=100, base=BASE):
    out = []
    for i in range(count):
        rng = SplitMix64(stream_seed(base, 2_000_000 + i))
        n = nmin + rng.below(nmax - nmin + 1)
        h = workloads.random_pauli_hamiltonian(n, terms, 4, stream_seed(base, 3_000_000 + i))
        out.append((n, layers, h))
    return out
