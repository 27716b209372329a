"""QAOA MaxCut on the device against the oracle's restatement of qaoa.rs.

* evaluate_qaoa_cost: exact mode gives amplitudes bit-identical to the
  oracle (f64) and the cost within 1e-12; fused mode within 1e-12 (f64) /
  1e-5 (f32);
* run_qaoa: the plain optimizer follows the oracle step for step (energy
  history, angles, best bitstring);
* the 26-node, p = 4 instance against the committed oracle fixture
  (tests/golden/qaoa_n26_p4.json), with 10^5 sampled cuts.
"""
import json
import math
import os

import numpy as np
import pytest

from ruqu_amd import workloads
from ruqu_amd.qaoa import Graph, QaoaConfig, QaoaInstance, cut_values, run_qaoa
from tests.helpers import orc

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "qaoa_n26_p4.json")


def graphs():
    out = []
    e = workloads.random_regular_graph(10, 3, 1)
    out.append(("3reg10", 10, e, [1.0] * len(e)))
    e, w = workloads.erdos_renyi_graph(9, 0.5, 2)
    out.append(("er9w", 9, e, w))
    e, w = workloads.erdos_renyi_graph(12, 0.5, 5)
    out.append(("er12w", 12, e, w))
    out.append(("tri", 3, [(0, 1), (1, 2), (0, 2)], [1.0, 1.0, 1.0]))
    return out


@pytest.mark.parametrize("name,n,edges,w", graphs())
def test_cost_exact_bit_identical(name, n, edges, w):
    print("qaoa exact", name, flush=True)
    g = Graph(n, [(i, j, x) for (i, j), x in zip(edges, w)])
    inst = QaoaInstance(g, "c64")
    gam, bet = [0.7, 1.9, 0.3], [0.4, 2.2, 1.1]
    cost = inst.cost(gam, bet, "exact")
    want_cost, want = orc.qaoa_cost(n, edges, w, gam, bet, want_state=True)
    assert np.array_equal(inst.state().state_vector(), want)
    assert abs(cost - want_cost) < 1e-12
    fused = inst.cost(gam, bet, "fused")
    assert abs(fused - want_cost) < 1e-12
    assert np.max(np.abs(inst.state().state_vector() - want)) < 1e-12


@pytest.mark.parametrize("name,n,edges,w", graphs())
def test_cost_f32(name, n, edges, w):
    g = Graph(n, [(i, j, x) for (i, j), x in zip(edges, w)])
    inst = QaoaInstance(g, "c32")
    gam, bet = [0.7, 1.9], [0.4, 2.2]
    want_cost, want = orc.qaoa_cost(n, edges, w, gam, bet, want_state=True)
    for mode in ("exact", "fused"):
        c = inst.cost(gam, bet, mode)
        assert abs(c - want_cost) < 1e-5 * max(1.0, len(edges))
        assert np.max(np.abs(inst.state().state_vector() - want)) < 1e-5
    # exact mode follows the f32 oracle bit for bit (QuantumStateF32 arithmetic)
    inst.cost(gam, bet, "exact")
    a32, _, _ = orc.run_circuit(n, orc.qaoa_circuit(n, edges, w, gam, bet), f32=True)
    assert np.array_equal(inst.state().state_vector(), a32)


def test_run_qaoa_matches_oracle():
    n = 8
    edges = workloads.random_regular_graph(n, 3, 4)
    w = [1.0] * len(edges)
    g = Graph.unweighted(n, edges)
    got = run_qaoa(QaoaConfig(g, 2, 12, 0.1), "c64", "exact")
    want = orc.run_qaoa(n, edges, w, 2, 12, 0.1)
    assert len(got.energy_history) == len(want["energy_history"])
    assert np.max(np.abs(got.energy_history - want["energy_history"])) < 1e-12
    assert np.max(np.abs(got.optimal_gammas - want["optimal_gammas"])) < 1e-12
    assert np.max(np.abs(got.optimal_betas - want["optimal_betas"])) < 1e-12
    assert got.best_index == want["best_index"]
    assert got.converged == want["converged"]
    fused = run_qaoa(QaoaConfig(g, 2, 12, 0.1), "c64", "fused")
    assert np.max(np.abs(fused.energy_history - want["energy_history"])) < 1e-10


def test_qaoa_26_nodes_p4_against_fixture():
    with open(GOLDEN) as f:
        gold = json.load(f)
    n, edges = gold["n"], [tuple(e) for e in gold["edges"]]
    print("qaoa n=26 p=4", flush=True)
    g = Graph.unweighted(n, edges)
    inst = QaoaInstance(g, "c64")
    cost = inst.cost(gold["gammas"], gold["betas"], "exact")
    assert abs(cost - gold["cost"]) < 1e-12
    st = inst.state()
    amps = st.state_vector()
    idx = np.array(gold["indices"])
    want = np.array([float.fromhex(r) for r in gold["amps_re_hex"]]) + 1j * np.array(
        [float.fromhex(r) for r in gold["amps_im_hex"]])
    assert np.array_equal(amps[idx], want)
    fused = inst.cost(gold["gammas"], gold["betas"], "fused")
    assert abs(fused - gold["cost"]) < 1e-12
    shots = 100_000
    draws = workloads.uniforms(workloads.shot_seed(3), shots * n).reshape(shots, n)
    bits, cuts, ties = inst.sample_cuts(draws)
    se = np.std(cuts) / math.sqrt(shots)
    print("  <H_C> %.6f sampled mean %.6f +- %.6f, best sampled cut %d, near ties %d" %
          (fused, cuts.mean(), se, cuts.max(), ties), flush=True)
    assert abs(cuts.mean() - fused) < 6 * se
    assert np.array_equal(cut_values(g, bits[:3]), [sum(1 for i, j in edges if b[i] != b[j]) for b in bits[:3]])
    # every one of the 10^5 shots against measure_all on a copy of the exact state
    ref = orc.measure_all_shots(amps, n, draws)
    mismatch = int(np.count_nonzero(np.any(bits != ref, axis=1)))
    print("  %d shots: %d differ from measure_all, near ties %d" % (shots, mismatch, ties), flush=True)
    assert mismatch <= ties


def test_sampled_bits_match_oracle_measure_all():
    n = 12
    edges = workloads.random_regular_graph(n, 3, 9)
    g = Graph.unweighted(n, edges)
    inst = QaoaInstance(g, "c64")
    inst.cost([0.4, 0.9], [0.7, 0.2], "exact")
    amps = inst.state().state_vector()
    shots = 300
    draws = workloads.uniforms(11, shots * n).reshape(shots, n)
    bits, _, ties = inst.sample_cuts(draws)
    bad = 0
    for s in range(shots):
        a = amps.copy()
        r, _ = orc.measure_all(a, n, draws[s])
        bad += int(list(bits[s]) != list(r))
    assert bad <= ties


def test_graph_validation():
    g = Graph(3)
    with pytest.raises(ValueError):
        g.add_edge(0, 5, 1.0)  # qaoa.rs:56-60 panics with "node index 5 out of range"
