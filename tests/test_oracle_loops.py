"""Pins the oracle's optimizer loops and ansatz builder against independent
computations, and shows by mutation that each check catches a plausible
mistake.

* run_qaoa (qaoa.rs:217-308): the expected cut comes from the textbook QAOA
  state, prod_l [exp(-i b_l X)^(x)n  diag(exp(-i g_l sum_e w_e z_i z_j))] |+>^n
  with z = +-1 (Rzz(2 g w) = exp(-i g w Z Z), Rx(2 b) = exp(-i b X)), built
  with numpy from first principles; the optimizer rule of qaoa.rs is written
  out below over that energy.  One and two iterations, a run that converges
  (the 1e-6 break), and the best bitstring (the LAST maximal index of the
  probabilities at the angles of the best cost, Iterator::max_by) are checked.
* run_vqe (vqe.rs:148-216): gradients are central finite differences of a
  dense energy (the parameter-shift rule is exact for Ry/Rz, whose generators
  have eigenvalues +-1/2), and the update is simultaneous.
* build_ansatz (vqe.rs:81-112): the gate list of a 3-qubit depth-2 ansatz is
  written out by hand, and its dense state is compared with the oracle's
  energy.

The mutation tests compile the oracle with -DORC_PLANT=k (one planted
mistake each, listed in oracle/ruqu_oracle.cpp) and require the matching
check to fail.
"""
import math
import os

import numpy as np
import pytest

from oracle import oracle as orc_mod
from ruqu_amd.gates import GateKind
from tests.helpers import PAULI, pauli_matrix

# ---------------------------------------------------------------------------
# Independent dense QAOA
# ---------------------------------------------------------------------------

N_Q = 6
_rng = np.random.default_rng(7)
EDGES = [(i, j) for i in range(N_Q) for j in range(i + 1, N_Q) if _rng.random() < 0.6]
WEIGHTS = [float(x) for x in np.round(_rng.uniform(0.2, 1.5, len(EDGES)), 3)]


def _bits(n):
    x = np.arange(1 << n)
    return 1 - 2 * ((x[:, None] >> np.arange(n)) & 1)  # z_q = +1 for bit 0, -1 for bit 1


def dense_qaoa_state(n, edges, w, gammas, betas):
    z = _bits(n)
    hzz = np.zeros(1 << n)
    for (i, j), wt in zip(edges, w):
        hzz += wt * z[:, i] * z[:, j]
    psi = np.full(1 << n, 2.0 ** (-n / 2), dtype=complex)
    for g, b in zip(gammas, betas):
        psi = psi * np.exp(-1j * g * hzz)
        rx = np.array([[math.cos(b), -1j * math.sin(b)], [-1j * math.sin(b), math.cos(b)]])
        t = psi.reshape([2] * n)
        for ax in range(n):
            t = np.moveaxis(np.tensordot(rx, t, axes=([1], [ax])), 0, ax)
        psi = t.reshape(-1)
    return psi


def dense_cut_table(n, edges, w):
    z = _bits(n)
    cut = np.zeros(1 << n)
    for (i, j), wt in zip(edges, w):
        cut += wt * 0.5 * (1 - z[:, i] * z[:, j])
    return cut


def dense_qaoa_cost(n, edges, w, gammas, betas):
    psi = dense_qaoa_state(n, edges, w, gammas, betas)
    return float(np.sum(np.abs(psi) ** 2 * dense_cut_table(n, edges, w)))


def numpy_run_qaoa(n, edges, w, p, max_iter, lr):
    """qaoa.rs:217-308 over the dense energy: gammas then betas updated in
    sequence with pi/2 shifts; stop when |C_k - C_(k-1)| < 1e-6."""
    g, b = [0.5] * p, [0.5] * p
    hist, angles, best, best_it, conv = [], [], -math.inf, -1, False
    for it in range(max_iter):
        c = dense_qaoa_cost(n, edges, w, g, b)
        hist.append(c)
        angles.append((list(g), list(b)))
        if c > best:
            best, best_it = c, it
        if it > 0 and abs(c - hist[it - 1]) < 1e-6:
            conv = True
            break
        s = math.pi / 2
        for i in range(p):
            gp, gm = list(g), list(g)
            gp[i] += s
            gm[i] -= s
            g[i] += lr * (dense_qaoa_cost(n, edges, w, gp, b) - dense_qaoa_cost(n, edges, w, gm, b)) / 2
        for i in range(p):
            bp, bm = list(b), list(b)
            bp[i] += s
            bm[i] -= s
            b[i] += lr * (dense_qaoa_cost(n, edges, w, g, bp) - dense_qaoa_cost(n, edges, w, g, bm)) / 2
    return {"hist": np.array(hist), "g": np.array(g), "b": np.array(b), "angles": angles, "best": best,
            "best_it": best_it, "converged": conv}


def check_run_qaoa_iterations(orc):
    """One and two full iterations, then 12 (p = 3, weighted graph)."""
    for iters in (1, 2, 3, 12):
        want = numpy_run_qaoa(N_Q, EDGES, WEIGHTS, 3, iters, 0.2)
        got = orc.run_qaoa(N_Q, EDGES, WEIGHTS, 3, iters, 0.2)
        assert len(got["energy_history"]) == len(want["hist"])
        assert np.max(np.abs(got["energy_history"] - want["hist"])) < 1e-11
        assert np.max(np.abs(got["optimal_gammas"] - want["g"])) < 1e-11
        # beta's shift gradient vanishes identically for MaxCut (a pi/2 shift of
        # every mixer angle is X on every qubit up to a phase, and the cut table
        # is invariant under flipping every bit), so beta stays at 0.5 up to rounding
        assert np.max(np.abs(got["optimal_betas"] - 0.5)) < 1e-12
        assert abs(got["best_cut_value"] - want["best"]) < 1e-11


def check_run_qaoa_convergence(orc):
    """A run that stops at the 1e-6 break (150 of 300 iterations)."""
    want = numpy_run_qaoa(N_Q, EDGES, WEIGHTS, 3, 300, 0.2)
    assert want["converged"] and len(want["hist"]) < 300
    got = orc.run_qaoa(N_Q, EDGES, WEIGHTS, 3, 300, 0.2)
    assert got["converged"] == want["converged"]
    assert len(got["energy_history"]) == len(want["hist"])
    assert np.max(np.abs(got["energy_history"] - want["hist"])) < 1e-9


def check_run_qaoa_argmax(orc):
    """best_bitstring: the last maximal probability at the angles of the best
    cost (qaoa.rs:238-259), taken before that iteration's update."""
    p, iters, lr = 2, 40, 0.1
    want = numpy_run_qaoa(N_Q, EDGES, WEIGHTS, p, iters, lr)
    got = orc.run_qaoa(N_Q, EDGES, WEIGHTS, p, iters, lr)
    k = want["best_it"]
    g, b = want["angles"][k]
    pr_dense = np.abs(dense_qaoa_state(N_Q, EDGES, WEIGHTS, g, b)) ** 2
    idx = got["best_index"]
    assert pr_dense[idx] >= pr_dense.max() - 1e-12
    assert got["best_bitstring"] == [bool((idx >> q) & 1) for q in range(N_Q)]
    # Ties: the state commutes with flipping every bit, so x and ~x are equally
    # likely and the reference's max_by picks the later index.  The oracle's own
    # probabilities at its angles of iteration k (a run stopped after k updates)
    # decide the tie bit for bit.
    at_k = orc.run_qaoa(N_Q, EDGES, WEIGHTS, p, k, lr) if k > 0 else None
    ga = at_k["optimal_gammas"] if at_k else np.full(p, 0.5)
    ba = at_k["optimal_betas"] if at_k else np.full(p, 0.5)
    _, amps = orc.qaoa_cost(N_Q, EDGES, WEIGHTS, ga, ba, want_state=True)
    pr = amps.real * amps.real + amps.imag * amps.imag
    ties = np.flatnonzero(pr == pr.max())
    assert len(ties) >= 2  # the tie the last-maximum rule decides
    assert idx == ties[-1]


# ---------------------------------------------------------------------------
# Independent dense VQE
# ---------------------------------------------------------------------------

VQE_N, VQE_DEPTH = 3, 2
VQE_TERMS = [(0.5, [(0, 3), (1, 3)]), (-0.7, [(1, 1), (2, 1)]), (0.3, [(0, 2), (2, 2)]), (0.9, [(2, 3)]),
             (-0.4, [(0, 1)]), (0.2, [(0, 1), (1, 2), (2, 3)]), (-0.25, [])]


def hand_written_ansatz(p):
    """vqe.rs:93-110 for 3 qubits, depth 2, written out gate by gate:
    (kind, qubit(s), parameter)."""
    R, Z, CX = GateKind.Ry, GateKind.Rz, GateKind.CNOT
    return [(R, 0, p[0]), (R, 1, p[1]), (R, 2, p[2]), (Z, 0, p[3]), (Z, 1, p[4]), (Z, 2, p[5]),
            (CX, (0, 1), None), (CX, (1, 2), None),
            (R, 0, p[6]), (R, 1, p[7]), (R, 2, p[8]), (Z, 0, p[9]), (Z, 1, p[10]), (Z, 2, p[11]),
            (CX, (0, 1), None), (CX, (1, 2), None)]


def _apply_dense(psi, n, q, m):
    t = psi.reshape([2] * n)
    ax = n - 1 - q  # C order: axis 0 is the highest index bit
    return np.moveaxis(np.tensordot(m, t, axes=([1], [ax])), 0, ax).reshape(-1)


def dense_ansatz_state(p):
    n = VQE_N
    psi = np.zeros(1 << n, dtype=complex)
    psi[0] = 1
    for kind, q, th in hand_written_ansatz(p):
        if kind == GateKind.CNOT:
            c, t = q
            idx = np.arange(1 << n)
            src = np.where((idx >> c) & 1, idx ^ (1 << t), idx)
            psi = psi[src]
        elif kind == GateKind.Ry:
            m = np.array([[math.cos(th / 2), -math.sin(th / 2)], [math.sin(th / 2), math.cos(th / 2)]])
            psi = _apply_dense(psi, n, q, m)
        else:
            m = np.diag([np.exp(-0.5j * th), np.exp(0.5j * th)])
            psi = _apply_dense(psi, n, q, m)
    return psi


def dense_energy(p):
    H = sum(c * pauli_matrix(VQE_N, ops) for c, ops in VQE_TERMS)
    psi = dense_ansatz_state(p)
    return float(np.real(np.conj(psi) @ H @ psi))


def numpy_run_vqe(max_iter, threshold, lr):
    """vqe.rs:148-216 with the gradient from central finite differences of the
    dense energy, all entries before any parameter moves."""
    npar = 2 * VQE_N * VQE_DEPTH
    prm = np.full(npar, 0.1)
    hist, best, best_p, conv = [], math.inf, prm.copy(), False
    for it in range(max_iter):
        e = dense_energy(prm)
        hist.append(e)
        if e < best:
            best, best_p = e, prm.copy()
        if it > 0 and abs(hist[it - 1] - e) < threshold:
            conv = True
            break
        h = 1e-5
        grad = np.zeros(npar)
        for i in range(npar):
            a, b = prm.copy(), prm.copy()
            a[i] += h
            b[i] -= h
            grad[i] = (dense_energy(a) - dense_energy(b)) / (2 * h)
        prm = prm - lr * grad
    return {"hist": np.array(hist), "best": best, "best_p": best_p, "converged": conv}


def _run_vqe(orc, max_iter, threshold, lr):
    return orc.run_vqe(VQE_N, VQE_DEPTH, VQE_TERMS, max_iter, threshold, lr)


def check_run_vqe_iterations(orc):
    for iters in (2, 3, 6):
        want = numpy_run_vqe(iters, 1e-12, 0.3)
        got = _run_vqe(orc, iters, 1e-12, 0.3)
        assert len(got["energy_history"]) == len(want["hist"])
        assert np.max(np.abs(got["energy_history"] - want["hist"])) < 1e-8


def check_run_vqe_best(orc):
    """optimal_parameters / optimal_energy: the minimum over the history, and
    the 1e-3 convergence break."""
    want = numpy_run_vqe(60, 1e-3, 0.3)
    got = _run_vqe(orc, 60, 1e-3, 0.3)
    assert want["converged"] and got["converged"]
    assert got["num_iterations"] == len(want["hist"])
    assert abs(got["optimal_energy"] - want["best"]) < 1e-8
    assert np.max(np.abs(got["optimal_parameters"] - want["best_p"])) < 1e-7


def check_build_ansatz(orc):
    prm = np.linspace(0.15, 1.9, 12)
    arr = orc.build_ansatz(VQE_N, VQE_DEPTH, prm)
    hand = hand_written_ansatz(prm)
    assert len(arr) == len(hand)
    for g, (kind, q, th) in zip(arr, hand):
        assert int(g["kind"]) == kind
        if kind == GateKind.CNOT:
            assert (int(g["q0"]), int(g["q1"])) == q
        else:
            assert int(g["q0"]) == q and float(g["theta"]) == th
    e = orc.energy(VQE_N, arr, VQE_TERMS)
    assert abs(e - dense_energy(prm)) < 1e-13


# ---------------------------------------------------------------------------
# Tests: the faithful oracle passes every check ...
# ---------------------------------------------------------------------------

CHECKS = {
    "qaoa_iterations": check_run_qaoa_iterations,
    "qaoa_convergence": check_run_qaoa_convergence,
    "qaoa_argmax": check_run_qaoa_argmax,
    "vqe_iterations": check_run_vqe_iterations,
    "vqe_best": check_run_vqe_best,
    "build_ansatz": check_build_ansatz,
}


def test_dense_textbook_qaoa_matches_oracle_cost():
    for gam, bet in ([0.5, 0.5, 0.5], [0.5, 0.5, 0.5]), ([0.3, -1.1, 2.0], [0.7, 0.2, -0.4]):
        got = orc_mod.qaoa_cost(N_Q, EDGES, WEIGHTS, gam, bet)
        assert abs(got - dense_qaoa_cost(N_Q, EDGES, WEIGHTS, gam, bet)) < 1e-12


def test_finite_difference_equals_parameter_shift():
    """The premise of the VQE checks: for Ry/Rz ansatz angles the pi/2 shift
    rule equals the derivative."""
    prm = np.linspace(-0.8, 1.3, 12)
    for i in (0, 5, 7, 11):
        a, b = prm.copy(), prm.copy()
        a[i] += math.pi / 2
        b[i] -= math.pi / 2
        shift = (dense_energy(a) - dense_energy(b)) / 2
        a, b = prm.copy(), prm.copy()
        a[i] += 1e-5
        b[i] -= 1e-5
        fd = (dense_energy(a) - dense_energy(b)) / 2e-5
        assert abs(shift - fd) < 1e-8


@pytest.mark.parametrize("name", list(CHECKS))
def test_oracle_passes(name):
    CHECKS[name](orc_mod)


# ---------------------------------------------------------------------------
# ... and every planted mistake fails the check meant to catch it.
# ---------------------------------------------------------------------------

PLANTS = {
    1: "qaoa_iterations",   # gamma step sign
    2: "qaoa_iterations",   # gamma gradients not in sequence
    3: "qaoa_iterations",   # shift on the wrong parameter
    4: "qaoa_argmax",       # first maximal index
    5: "qaoa_convergence",  # convergence against the first cost
    6: "qaoa_argmax",       # bitstring at the updated angles
    7: "vqe_iterations",    # descent sign
    8: "vqe_iterations",    # one-by-one update
    9: "vqe_best",          # best tracked at the maximum
    10: "build_ansatz",     # CNOT chain reversed
    11: "build_ansatz",     # Rz before Ry
    12: "build_ansatz",     # parameter index restarts per layer
}


@pytest.fixture(scope="module")
def mutant_dir(tmp_path_factory):
    return str(tmp_path_factory.mktemp("oracle_plants"))


@pytest.mark.parametrize("k", sorted(PLANTS))
def test_planted_mistake_is_caught(k, mutant_dir):
    if not os.path.exists(os.path.join(orc_mod.HERE, "_build", "simd.o")):
        orc_mod.build()
    path = orc_mod.build_planted(k, mutant_dir)
    with orc_mod.planted(path):
        with pytest.raises(AssertionError):
            CHECKS[PLANTS[k]](orc_mod)
