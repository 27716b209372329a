"""Pins the oracle against what the mathematics fixes, independently of the
oracle's own formulas:

* dense Kronecker-product operators for every gate variant and index order
  (a transposed operand, a wrong qubit bit or a sign error fails here);
* expectation values against dense Pauli matrices, including Y phases and
  repeated qubits;
* measurement probabilities / collapse against numpy sums;
* the QAOA p=1 MaxCut closed form (Wang, Hadfield, Jiang, Rieffel 2018);
* the exact ground energy of the coded H2 Hamiltonian (eigvalsh);
* the simd.rs restatement bit-identical to the scalar restatement;
* fuse_gates equal to the unfused circuit; run_shots equal to measure_all
  on copies given the same draws;
* the reference's validation rules (state.rs:31-96, :179-184, :441-449).
"""
import math

import numpy as np
import pytest

from ruqu_amd.gates import Gate, GateKind, QuantumCircuit
from ruqu_amd import workloads
from tests.helpers import (ALL_1Q, ALL_2Q, embed_1q, embed_2q, max_abs, mixed_circuit, orc, pauli_matrix,
                           random_state, random_unitary)


def dense_unitary_of(n, arr):
    """Product of embedded gate matrices, computed from closed forms here."""
    U = np.eye(1 << n, dtype=np.complex128)
    for g in arr:
        k = int(g["kind"])
        th = float(g["theta"])
        q0, q1 = int(g["q0"]), int(g["q1"])
        if k == GateKind.Barrier:
            continue
        if k in (GateKind.CNOT, GateKind.CZ, GateKind.SWAP, GateKind.Rzz):
            if k == GateKind.CNOT:
                m = np.array([[1, 0, 0, 0], [0, 1, 0, 0], [0, 0, 0, 1], [0, 0, 1, 0]], dtype=complex)
            elif k == GateKind.CZ:
                m = np.diag([1, 1, 1, -1]).astype(complex)
            elif k == GateKind.SWAP:
                m = np.array([[1, 0, 0, 0], [0, 0, 1, 0], [0, 1, 0, 0], [0, 0, 0, 1]], dtype=complex)
            else:
                m = np.diag(np.exp(-0.5j * th * np.array([1, -1, -1, 1])))
            U = embed_2q(n, q0, q1, m) @ U
            continue
        c, s = math.cos(th / 2), math.sin(th / 2)
        m = {
            GateKind.H: np.array([[1, 1], [1, -1]]) / math.sqrt(2),
            GateKind.X: np.array([[0, 1], [1, 0]]),
            GateKind.Y: np.array([[0, -1j], [1j, 0]]),
            GateKind.Z: np.diag([1, -1]),
            GateKind.S: np.diag([1, 1j]),
            GateKind.Sdg: np.diag([1, -1j]),
            GateKind.T: np.diag([1, np.exp(1j * math.pi / 4)]),
            GateKind.Tdg: np.diag([1, np.exp(-1j * math.pi / 4)]),
            GateKind.Rx: np.array([[c, -1j * s], [-1j * s, c]]),
            GateKind.Ry: np.array([[c, -s], [s, c]]),
            GateKind.Rz: np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)]),
            GateKind.Phase: np.diag([1, np.exp(1j * th)]),
        }.get(k)
        if k == GateKind.Unitary1Q:
            m = (g["u"][0::2] + 1j * g["u"][1::2]).reshape(2, 2)
        U = embed_1q(n, q0, np.asarray(m, dtype=complex)) @ U
    return U


@pytest.mark.parametrize("n,seed", [(3, 1), (5, 2), (7, 3)])
def test_oracle_matches_dense_operator(n, seed):
    c = mixed_circuit(n, 4, seed)
    arr = c.to_array()
    amps, _, cnt = orc.run_circuit(n, arr)
    want = dense_unitary_of(n, arr)[:, 0]
    assert max_abs(amps, want) < 1e-13
    assert cnt == len(arr)


@pytest.mark.parametrize("kind", ALL_1Q)
def test_every_1q_gate_every_qubit(kind):
    n = 4
    for q in range(n):
        psi = random_state(n, 10 + q)
        g = QuantumCircuit(n).add_gate(Gate(kind, q, 0, 0.77)).to_array()
        amps = psi.copy()
        orc.apply_gate(amps, n, g[0])
        assert max_abs(amps, dense_unitary_of(n, g) @ psi) < 1e-14


@pytest.mark.parametrize("kind", ALL_2Q)
def test_every_2q_gate_every_pair(kind):
    n = 4
    for q1 in range(n):
        for q2 in range(n):
            if q1 == q2:
                continue
            psi = random_state(n, q1 * 7 + q2)
            g = QuantumCircuit(n).add_gate(Gate(kind, q1, q2, 0.61)).to_array()
            amps = psi.copy()
            orc.apply_gate(amps, n, g[0])
            assert max_abs(amps, dense_unitary_of(n, g) @ psi) < 1e-14


def test_f32_tracks_f64():
    n = 8
    arr = mixed_circuit(n, 6, 5).to_array()
    a64, _, _ = orc.run_circuit(n, arr)
    a32, _, _ = orc.run_circuit(n, arr, f32=True)
    assert a32.dtype == np.complex64
    assert max_abs(a32, a64) < 1e-5


@pytest.mark.parametrize("seed", range(4))
def test_expectation_against_dense_pauli(seed):
    n = 5
    rng = np.random.default_rng(seed)
    psi = random_state(n, 100 + seed)
    for _ in range(20):
        w = int(rng.integers(0, 5))
        qs = rng.choice(n, w, replace=False)
        ops = [(int(q), int(rng.integers(4))) for q in qs]
        P = pauli_matrix(n, ops)
        want = float(np.real(np.vdot(psi, P @ psi)))
        assert abs(orc.expectation(psi, n, ops) - want) < 1e-13


def test_expectation_repeated_qubit_quirk():
    """state.rs:370-392 reads every factor's bit from the input index i (not
    from the partially flipped j), so a repeated qubit does not compose like
    the operator product: Y_q Y_q contributes i*i = -1 instead of the
    identity, and Z_q Y_q contributes i.  The oracle keeps the reference's
    behaviour."""
    n = 4
    psi = random_state(n, 5)
    x3 = float(np.real(np.vdot(psi, pauli_matrix(n, [(3, 1)]) @ psi)))
    assert abs(orc.expectation(psi, n, [(2, 2), (2, 2), (3, 1)]) - (-x3)) < 1e-13
    y2 = pauli_matrix(n, [(2, 2)])
    z2 = pauli_matrix(n, [(2, 3)])
    # Z then Y on qubit 2: phase i * (-1)^b * (-1)^b = i, flip of bit 2
    flip = pauli_matrix(n, [(2, 1)])
    want = float(np.real(np.vdot(psi, 1j * flip @ psi)))
    assert abs(orc.expectation(psi, n, [(2, 3), (2, 2)]) - want) < 1e-13
    assert not np.allclose(z2 @ y2, 1j * flip)  # the true product differs


def test_hamiltonian_sum():
    n = 4
    psi = random_state(n, 7)
    terms = workloads.random_pauli_hamiltonian(n, 12, 3, 99)
    want = sum(c * float(np.real(np.vdot(psi, pauli_matrix(n, ops) @ psi))) for c, ops in terms)
    assert abs(orc.expectation_hamiltonian(psi, n, terms) - want) < 1e-12


def test_measure_probability_and_collapse():
    n = 5
    psi = random_state(n, 3)
    for q in range(n):
        p1 = float(np.sum(np.abs(psi[[i for i in range(1 << n) if i >> q & 1]]) ** 2))
        assert abs(orc.probability_of_qubit(psi, n, q) - p1) < 1e-15
        for draw in (0.0, 1 - p1 - 1e-9, 1 - p1 + 1e-9, 0.999):
            a = psi.copy()
            res, prob = orc.measure(a, n, q, draw)
            assert res == (draw >= 1 - p1 if abs(draw - (1 - p1)) > 1e-12 else res)
            keep = np.array([(i >> q & 1) == int(res) for i in range(1 << n)])
            assert np.allclose(a[~keep], 0)
            assert abs(np.sum(np.abs(a) ** 2) - 1) < 1e-13
            assert abs(prob - (p1 if res else 1 - p1)) < 1e-14


def qaoa_p1_closed_form(n, edges, gamma, beta):
    """Expected cut of p=1 QAOA on an unweighted graph (Wang et al. 2018,
    Theorem 1) for this circuit's convention: Rzz(2*gamma) = exp(-i gamma Z Z),
    Rx(2*beta) = exp(-i beta X).  In the paper's exp(-i g' C) exp(-i b' B) with
    C = sum (1 - Z Z)/2 this is g' = -2 gamma, b' = beta."""
    g, b = -2.0 * gamma, beta
    adj = {v: set() for v in range(n)}
    for u, v in edges:
        adj[u].add(v)
        adj[v].add(u)
    total = 0.0
    for u, v in edges:
        d = len(adj[u]) - 1
        e = len(adj[v]) - 1
        f = len(adj[u] & adj[v])
        term = 0.5 + 0.25 * math.sin(4 * b) * math.sin(g) * (math.cos(g) ** d + math.cos(g) ** e)
        term -= 0.25 * math.sin(2 * b) ** 2 * math.cos(g) ** (d + e - 2 * f) * (1 - math.cos(2 * g) ** f)
        total += term
    return total


@pytest.mark.parametrize("edges,n", [
    ([(0, 1)], 2),
    ([(0, 1), (1, 2), (0, 2)], 3),
    ([(0, 1), (1, 2), (2, 3), (0, 3)], 4),
    ([(0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)], 4),
])
def test_qaoa_p1_closed_form(edges, n):
    for gamma, beta in [(0.3, 0.2), (0.5, 0.5), (1.1, 2.3), (-0.7, 0.9)]:
        got = orc.qaoa_cost(n, edges, [1.0] * len(edges), [gamma], [beta])
        assert abs(got - qaoa_p1_closed_form(n, edges, gamma, beta)) < 1e-12


def test_qaoa_cost_matches_dense_expectation():
    n = 6
    edges, w = workloads.erdos_renyi_graph(n, 0.5, 11)
    gam, bet = [0.4, 1.1], [0.3, 0.8]
    cost, amps = orc.qaoa_cost(n, edges, w, gam, bet, want_state=True)
    U = dense_unitary_of(n, orc.qaoa_circuit(n, edges, w, gam, bet))
    psi = U[:, 0]
    assert max_abs(amps, psi) < 1e-13
    probs = np.abs(psi) ** 2
    cuts = np.array([sum(wt for (i, j), wt in zip(edges, w) if (x >> i & 1) != (x >> j & 1)) for x in range(1 << n)])
    assert abs(cost - float(probs @ cuts)) < 1e-12


def test_run_qaoa_improves_and_bitstring_is_argmax():
    """best_cut_value is the largest cost of the history, and best_bitstring a
    most probable outcome of the textbook QAOA state at the angles where that
    cost was reached (qaoa.rs:238-259).  The full loop is pinned step by step
    in tests/test_oracle_loops.py."""
    from tests.test_oracle_loops import dense_qaoa_state
    n = 5
    edges = [(0, 1), (1, 2), (2, 3), (3, 4), (4, 0)]
    w = [1.0] * 5
    r = orc.run_qaoa(n, edges, w, 2, 15, 0.1)
    hist = r["energy_history"]
    assert len(hist) <= 15
    assert r["best_cut_value"] == hist.max()
    k = int(np.argmax(hist))  # strict `cost > best`: the first iteration reaching the maximum
    at_k = orc.run_qaoa(n, edges, w, 2, k, 0.1) if k > 0 else None
    g = at_k["optimal_gammas"] if at_k else [0.5, 0.5]
    b = at_k["optimal_betas"] if at_k else [0.5, 0.5]
    pr = np.abs(dense_qaoa_state(n, edges, w, g, b)) ** 2
    idx = r["best_index"]
    assert pr[idx] >= pr.max() - 1e-12
    assert r["best_bitstring"] == [bool((idx >> q) & 1) for q in range(n)]


def h2_terms():  # vqe.rs:231-267
    return [(-1.0523, []), (0.3979, [(1, 3)]), (-0.3979, [(0, 3)]), (-0.0112, [(0, 3), (1, 3)]),
            (0.1809, [(0, 1), (1, 1)])]


def test_h2_ground_energy_and_vqe():
    H = sum(c * pauli_matrix(2, ops) for c, ops in h2_terms())
    e0 = float(np.linalg.eigvalsh(H)[0])
    assert abs(e0 - (-1.857202)) < 1e-5
    r = orc.run_vqe(2, 2, h2_terms(), 100, 1e-8, 0.1)
    assert r["optimal_energy"] >= e0 - 1e-9
    assert r["optimal_energy"] < -0.8  # test_algorithms.rs:341


def test_ansatz_matches_build_rule():
    params = np.linspace(0.1, 1.2, 2 * 3 * 2)
    arr = orc.build_ansatz(3, 2, params)
    kinds = [int(k) for k in arr["kind"]]
    assert kinds == [GateKind.Ry] * 3 + [GateKind.Rz] * 3 + [GateKind.CNOT] * 2 + [GateKind.Ry] * 3 + \
        [GateKind.Rz] * 3 + [GateKind.CNOT] * 2


def test_fuse_gates_preserves_state():  # optimizer.rs:46-94
    n = 4
    c = QuantumCircuit(n)
    for k in range(30):
        c.rx(k % 2, 0.1 * k).ry(k % 2, 0.2 * k).rz(k % 2, 0.3).cnot(2, 3) if k % 5 == 0 else c.h(k % 4).rz(k % 4, 0.5)
    arr = c.to_array()
    fused = orc.fuse_gates(arr)
    assert len(fused) < len(arr)
    assert np.any(fused["kind"] == GateKind.Unitary1Q)
    a, _, _ = orc.run_circuit(n, arr)
    b, _, _ = orc.run_circuit(n, fused)
    assert max_abs(a, b) < 1e-13


@pytest.mark.parametrize("n", [6, 17])
def test_simd_restatement_bit_identical(n):
    """simd.rs *_best: at >= 65 536 amplitudes the rayon path (scalar
    butterflies in chunks) is bit-identical to the scalar kernel.  The AVX2
    kernel (used below that size) builds its sign mask with
    _mm256_set_pd(-1, 1, -1, 1) (simd.rs:135), i.e. [+1, -1, +1, -1] in
    element order, so it multiplies by conj(m): bit-identical to the scalar
    kernel applied to the conjugated matrix.  The restatement keeps that."""
    rng = np.random.default_rng(n)
    psi = random_state(n, 42)
    for q in range(n):
        m = random_unitary(2, q)
        a = psi.copy()
        orc.apply_1q(a, n, q, m)
        ac = psi.copy()
        orc.apply_1q(ac, n, q, np.conj(m))
        b = psi.copy()
        orc.simd_apply_1q_best(b, n, q, m, 4)
        # qubit 0 has step 1: the AVX2 kernel falls back to scalar (simd.rs:176-180)
        assert np.array_equal(b, a if (n >= 17 or q == 0) else ac)
        if q >= 1:
            c = psi.copy()
            orc.simd_apply_1q_avx2(c, n, q, m)
            assert np.array_equal(c, ac)
    for _ in range(4):
        q1, q2 = [int(x) for x in rng.choice(n, 2, replace=False)]
        m = random_unitary(4, q1 + 10 * q2)
        a = psi.copy()
        b = psi.copy()
        orc.apply_2q(a, n, q1, q2, m)
        orc.simd_apply_2q_best(b, n, q1, q2, m, 4)
        assert np.array_equal(a, b)


def test_run_shots_equals_measure_all_on_copies():  # simulator.rs:107-164
    n = 6
    arr = mixed_circuit(n, 3, 8).to_array()
    shots = 40
    draws = workloads.uniforms(5, shots * n).reshape(shots, n)
    bits = orc.run_shots(n, arr, shots, draws)
    final, _, _ = orc.run_circuit(n, arr)
    for s in range(shots):
        a = final.copy()
        res, _ = orc.measure_all(a, n, draws[s])
        assert list(bits[s]) == list(res)


@pytest.mark.parametrize("dtype", [np.complex128, np.complex64])
def test_measure_all_shots_equals_measure_all_on_copies(dtype):
    """The memoised many-shot measure_all gives every shot the bits that
    measure_all on its own copy of the state gives (state.rs:329-335,
    mixed_precision.rs:473-515 for complex-f32), including shots whose draws
    sit exactly on a threshold."""
    n, shots = 9, 400
    psi = random_state(n, 21, dtype)
    draws = workloads.uniforms(13, shots * n).reshape(shots, n)
    a = psi.copy()
    _, p0 = orc.measure(a, n, 0, 0.0)
    draws[:5, 0] = p0  # draw == p0 measures 1 (state.rs:302)
    bits = orc.measure_all_shots(psi, n, draws)
    for s in range(shots):
        a = psi.copy()
        r, _ = orc.measure_all(a, n, draws[s])
        assert list(bits[s]) == [int(x) for x in r], s
    assert np.all(bits[:5, 0] == 1)


def test_run_shots_with_midcircuit_measure():
    n = 3
    c = QuantumCircuit(n).h(0).cnot(0, 1).measure(0).h(2).measure(2)
    shots = 20
    draws = workloads.uniforms(9, shots * 2).reshape(shots, 2)
    bits = orc.run_shots(n, c.to_array(), shots, draws)
    for s in range(shots):
        assert bits[s][0] == (draws[s][0] >= 0.5 - 1e-12)
        assert bits[s][1] == 0  # never measured: stays false (simulator.rs:139-145)


def test_validation_rules():
    assert orc.lib().orc_state_check(0, 0, 0) == 5      # new(0): CircuitError
    assert orc.lib().orc_state_check(0, 0, 1) == 0      # new_with_seed(0) accepted
    assert orc.lib().orc_state_check(33, 0, 1) == 1     # > MAX_QUBITS
    assert orc.lib().orc_state_check(0, 1, 1) == 5      # f32 rejects 0 qubits
    assert orc.lib().orc_state_check(34, 1, 1) == 1     # > MAX_QUBITS_F32
    amps = np.zeros(4, dtype=np.complex128)
    amps[0] = 1
    rc, _ = orc.apply_gate(amps, 2, QuantumCircuit(2).h(5).to_array()[0])
    assert rc == 2
    rc, _ = orc.apply_gate(amps, 2, QuantumCircuit(2).cnot(1, 1).to_array()[0])
    assert rc == 5


def test_workload_generators_deterministic():
    a = workloads.random_circuit(12, 3, 1).to_array()
    b = workloads.random_circuit(12, 3, 1).to_array()
    assert np.array_equal(a, b)
    assert len(a) == 3 * (3 * 12 + 6)
    e = workloads.random_regular_graph(26, 3, 3)
    assert len(e) == 39
    deg = np.bincount(np.array(e).ravel(), minlength=26)
    assert np.all(deg == 3)
    u = workloads.uniforms(7, 5)
    sm = workloads.SplitMix64(7)
    assert np.array_equal(u, [sm.uniform() for _ in range(5)])
