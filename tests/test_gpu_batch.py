"""Batched QAOA / VQE on the device against the oracle.

The batch runs complex-f32 states with FP64 reductions, so energies are
compared with the FP64 oracle at |dE| <= 1e-5 * (1 + sum |weights or
coefficients|) (the f32 amplitude bar scaled by the observable's norm).
Instances cover both kernel paths: <= 13 qubits (state in LDS) and 14..20
qubits (state streamed through LDS tiles from HBM).
"""
import math

import numpy as np
import pytest

from ruqu_amd import workloads
from ruqu_amd.batch import Batch
from tests.helpers import orc

pytestmark = pytest.mark.gpu


def qaoa_cases():
    out = []
    for k, (n, p) in enumerate([(5, 1), (8, 2), (10, 3), (13, 2), (14, 1), (16, 2), (20, 1)]):
        e, w = workloads.erdos_renyi_graph(n, 0.5, 100 + k)
        rng = np.random.default_rng(k)
        out.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
    return out


def vqe_cases():
    out = []
    for k, n in enumerate([4, 12, 13, 14, 15]):
        terms = workloads.random_pauli_hamiltonian(n, 30, 4, 200 + k)
        rng = np.random.default_rng(50 + k)
        out.append((n, terms, list(rng.uniform(-1, 1, n * 4))))
    return out


def test_batch_energies_match_oracle():
    b = Batch()
    qc, vc = qaoa_cases(), vqe_cases()
    for n, e, w, p, g, be in qc:
        b.add_qaoa(n, e, w, p, g, be)
    for n, terms, prm in vc:
        b.add_vqe(n, 4, terms, prm)
    en, amax = b.evaluate()
    print("batch evaluate %.3f ms" % b.last_ms(), flush=True)
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want = orc.qaoa_cost(n, e, w, g, be)
        tol = 1e-5 * (1 + sum(w))
        print("  qaoa n=%d p=%d E=%.8f oracle %.8f diff %.2e" % (n, p, en[k], want, en[k] - want), flush=True)
        assert abs(en[k] - want) <= tol
        if n <= 14:
            _, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
            pr = np.abs(amps) ** 2
            assert pr[amax[k]] >= pr.max() * (1 - 1e-4)
    for k, (n, terms, prm) in enumerate(vc):
        gates = workloads.ry_cz_ansatz(n, 4, prm).to_array()
        want = orc.energy(n, gates, terms)
        tol = 1e-5 * (1 + sum(abs(c) for c, _ in terms))
        got = en[len(qc) + k]
        print("  vqe n=%d E=%.8f oracle %.8f diff %.2e" % (n, got, want, got - want), flush=True)
        assert abs(got - want) <= tol


def oracle_vqe_rycz(n, layers, terms, max_iter, thr, lr):
    """vqe.rs:148-216 with the Ry-CZ ansatz, energies from the oracle."""
    params = np.full(n * layers, 0.1)
    hist, best = [], float("inf")
    for it in range(max_iter):
        e = orc.energy(n, workloads.ry_cz_ansatz(n, layers, params).to_array(), terms)
        hist.append(e)
        best = min(best, e)
        if it > 0 and abs(hist[it - 1] - e) < thr:
            break
        grad = np.zeros_like(params)
        for i in range(len(params)):
            pp, pm = params.copy(), params.copy()
            pp[i] += math.pi / 2
            pm[i] -= math.pi / 2
            grad[i] = (orc.energy(n, workloads.ry_cz_ansatz(n, layers, pp).to_array(), terms)
                       - orc.energy(n, workloads.ry_cz_ansatz(n, layers, pm).to_array(), terms)) / 2.0
        params = params - lr * grad
    return np.array(hist), params, best


@pytest.mark.parametrize("use_graph", [True, False])
def test_batch_optimizer_follows_reference_loops(use_graph):
    b = Batch()
    qa = []
    for k, n in enumerate([6, 9]):
        e = workloads.random_regular_graph(n if n % 2 == 0 else n + 1, 3, 7 + k)
        n2 = n if n % 2 == 0 else n + 1
        qa.append((n2, e, [1.0] * len(e), 2))
        b.add_qaoa(n2, e, [1.0] * len(e), 2)
    terms = workloads.random_pauli_hamiltonian(4, 10, 2, 3)
    vid = b.add_vqe(4, 2, terms)
    iters = 4
    ms = b.optimize(iters, qaoa_lr=0.1, vqe_lr=0.1, vqe_threshold=1e-9, use_graph=use_graph)
    print("batch optimize %d iterations %.3f ms (graph=%s)" % (iters, ms, use_graph), flush=True)
    for k, (n, e, w, p) in enumerate(qa):
        want = orc.run_qaoa(n, e, w, p, iters, 0.1)
        got = b.result(k)
        scale = 1e-5 * (1 + sum(w)) * 10
        assert len(got["energy_history"]) == len(want["energy_history"])
        assert np.max(np.abs(got["energy_history"] - want["energy_history"])) < scale
        assert np.max(np.abs(got["optimal_gammas"] - want["optimal_gammas"])) < 1e-3
        assert np.max(np.abs(got["optimal_betas"] - want["optimal_betas"])) < 1e-3
        assert orc.cut_value(e, w, got["best_bitstring"]) == orc.cut_value(e, w, want["best_bitstring"])
    hist, params, best = oracle_vqe_rycz(4, 2, terms, iters, 1e-9, 0.1)
    got = b.result(vid)
    assert len(got["energy_history"]) == len(hist)
    assert np.max(np.abs(got["energy_history"] - hist)) < 1e-4
    assert np.max(np.abs(got["params"] - params)) < 1e-3
    assert abs(got["best"] - best) < 1e-4


@pytest.mark.parametrize("exact", [True, False])
@pytest.mark.parametrize("tq", [9, 11, 12, 13])
def test_batch_tile_widths_match_oracle(tq, exact):
    """Wide states at every tile width.  Exact tiles include tiling B with
    0-3 low bits (n close to 2*tq: tiles sharing 128-byte lines, the
    XCD-swizzled order); widened tiles put the larger states on wider tiles,
    so every pass launches once per tile width."""
    sizes = [(14, 3), (16, 1), (17, 2), (18, 1)] + ([] if tq == 9 else [(20, 2)])
    b = Batch()
    qc = []
    for k, (n, p) in enumerate(sizes):
        e, w = workloads.erdos_renyi_graph(n, 0.4, 600 + k)
        rng = np.random.default_rng(80 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    terms = workloads.random_pauli_hamiltonian(15, 25, 4, 700)
    prm = list(np.random.default_rng(90).uniform(-1, 1, 15 * 3))
    b.add_vqe(15, 3, terms, prm)
    b.set_tile(tq, exact=exact)
    en, amax = b.evaluate()
    print("tile %d exact %s: evaluate %.3f ms" % (tq, exact, b.last_ms()), flush=True)
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
        assert abs(en[k] - want) <= 1e-5 * (1 + sum(w)), (tq, n, en[k], want)
        pr = np.abs(amps) ** 2
        assert pr[amax[k]] >= pr.max() * (1 - 1e-4)
    want = orc.energy(15, workloads.ry_cz_ansatz(15, 3, prm).to_array(), terms)
    assert abs(en[-1] - want) <= 1e-5 * (1 + sum(abs(c) for c, _ in terms))


def test_batch_widened_tiles_vqe_and_qaoa_match_oracle():
    """A 21-qubit VQE state (its Pauli energies folded over 13-qubit tiles)
    and a 22-qubit QAOA state widened to 13-qubit tiles beside a 16-qubit
    state on 12-qubit tiles: every pass launches once per width."""
    b = Batch()
    qc = []
    for k, (n, p) in enumerate([(22, 2), (16, 3)]):
        e, w = workloads.erdos_renyi_graph(n, 0.3, 900 + k)
        rng = np.random.default_rng(95 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    terms = workloads.random_pauli_hamiltonian(21, 12, 4, 901)
    prm = list(np.random.default_rng(96).uniform(-1, 1, 21 * 2))
    b.add_vqe(21, 2, terms, prm)
    b.set_tile(12)
    en, amax = b.evaluate()
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
        assert abs(en[k] - want) <= 1e-5 * (1 + sum(w)), (n, en[k], want)
        pr = np.abs(amps) ** 2
        assert pr[amax[k]] >= pr.max() * (1 - 1e-4)
    want = orc.energy(21, workloads.ry_cz_ansatz(21, 2, prm).to_array(), terms)
    assert abs(en[-1] - want) <= 1e-5 * (1 + sum(abs(c) for c, _ in terms)), (en[-1], want)


def test_batch_24_qubits_tile_widths_agree():
    e, w = workloads.erdos_renyi_graph(24, 0.3, 800)
    b = Batch()
    b.add_qaoa(24, e, w, 2, [0.4, 0.9], [0.3, 0.6])
    got = {}
    for tq in (12, 13):
        b.set_tile(tq, exact=True)
        got[tq] = b.evaluate()
    assert abs(got[12][0][0] - got[13][0][0]) <= 1e-5 * (1 + sum(w))
    b.set_tile(12)  # widened: 24 qubits run on 13-qubit tiles, re-planned on a ready batch
    assert abs(b.evaluate()[0][0] - got[13][0][0]) <= 1e-5 * (1 + sum(w))


def test_set_all_params_equals_per_instance_updates():
    qc, vc = qaoa_cases(), vqe_cases()

    def build():
        b = Batch()
        for n, e, w, p, g, be in qc:
            b.add_qaoa(n, e, w, p, g, be)
        for n, terms, prm in vc:
            b.add_vqe(n, 4, terms, prm)
        return b

    rng = np.random.default_rng(123)
    new = [rng.uniform(0, math.pi, 2 * p) for *_, p, _, _ in qc] + [rng.uniform(-1, 1, n * 4) for n, _, _ in vc]
    one, bulk = build(), build()
    one.evaluate()
    bulk.evaluate()
    for i, prm in enumerate(new):
        one.set_params(i, prm)
    bulk.set_all_params(new)
    e1, a1 = one.evaluate()
    e2, a2 = bulk.evaluate()
    assert np.array_equal(e1, e2) and np.array_equal(a1, a2)
    with pytest.raises(Exception):
        bulk.set_all_params(new[:-1])
