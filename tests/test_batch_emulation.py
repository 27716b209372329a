"""The batched evaluation kernel's source run on the CPU (host threads, one
per work-item, real barriers; tests/emu) over the records that batch.cpp
builds for the device, against the oracle.  Covers the kernel's persistent
work loop, both state placements (<= 13 qubits in LDS, more streamed through
LDS tiles from an HBM slot), the cut tables, the Pauli-term encoding and the
host packing, without a GPU.  Tolerance as on the device: complex-f32 states
with FP64 reductions, |dE| <= 1e-5 * (1 + sum |weights or coefficients|)."""
import math

import numpy as np
import pytest

from ruqu_amd import workloads
from ruqu_amd.batch import Batch
from tests.emu import batch_emu
from tests.helpers import orc


def test_batch_kernel_on_cpu_matches_oracle():
    b = Batch()
    qc = []
    for k, (n, p) in enumerate([(5, 1), (9, 2), (14, 1)]):
        e, w = workloads.erdos_renyi_graph(n, 0.5, 300 + k)
        rng = np.random.default_rng(30 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    vc = []
    for k, n in enumerate([4, 14]):
        terms = workloads.random_pauli_hamiltonian(n, 12, 4, 400 + k)
        prm = list(np.random.default_rng(60 + k).uniform(-1, 1, n * 2))
        vc.append((n, terms, prm))
        b.add_vqe(n, 2, terms, prm)
    en, am = batch_emu.evaluate(b, grid=3)
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
        assert abs(en[k] - want) <= 1e-5 * (1 + sum(w)), (n, en[k], want)
        pr = np.abs(amps) ** 2
        assert pr[am[k]] >= pr.max() * (1 - 1e-4)
    for k, (n, terms, prm) in enumerate(vc):
        want = orc.energy(n, workloads.ry_cz_ansatz(n, 2, prm).to_array(), terms)
        got = en[len(qc) + k]
        assert abs(got - want) <= 1e-5 * (1 + sum(abs(c) for c, _ in terms)), (n, got, want)


def test_batch_plan_order_and_layout():
    b = Batch()
    e, w = workloads.erdos_renyi_graph(16, 0.5, 5)
    b.add_qaoa(16, e, w, 2)
    e2, w2 = workloads.erdos_renyi_graph(6, 0.5, 6)
    b.add_qaoa(6, e2, w2, 1)
    b.add_vqe(12, 3, workloads.random_pauli_hamiltonian(12, 5, 3, 7))
    e3, w3 = workloads.erdos_renyi_graph(14, 0.5, 8)
    b.add_qaoa(14, e3, w3, 3)
    d = batch_emu.data(b)
    ins = d["insts"]
    assert list(ins["n"]) == [16, 6, 12, 14]
    assert list(ins["nparams"]) == [4, 2, 36, 6]
    assert list(ins["ctab_off"]) == [0, 1 << 16, -1, (1 << 16) + (1 << 6)]  # a C(i) table per QAOA instance
    assert np.isclose(ins[0]["wsum"], sum(w))
    assert len(d["params"]) == 4 + 2 + 36 + 6
    assert np.all(d["params"][:6] == 0.5) and np.all(d["params"][6:42] == 0.1) and np.all(d["params"][42:] == 0.5)
    r = batch_emu.plan(b, -1)
    tq = r["tq"]
    assert tq == 12 and r["ctab_total"] == (1 << 16) + (1 << 6) + (1 << 14)
    assert list(r["small"]["inst"]) == [2, 1]  # largest state first
    assert list(r["wide"]["inst"]) == [3, 0]  # most layers first: layer l launches a prefix
    assert list(r["wtile0"]) == [0, 1 << (14 - tq)]
    assert list(r["wstate"]) == [0, 1 << 14]
    assert r["ntiles"] == (1 << (14 - tq)) + (1 << (16 - tq)) and len(r["wmap"]) == r["ntiles"]
    # p + 1 passes per item: both items in passes 0-2, the 14-qubit p=3 item alone in pass 3
    assert list(r["pass_grid"]) == [r["ntiles"], r["ntiles"], r["ntiles"], 1 << (14 - tq)]
    assert r["nvtiles"] == 0
    assert r["state_amps"] == (1 << 14) + (1 << 16)
    # round 1 of an optimizer iteration: the two gamma_0-shifted circuits of every QAOA instance
    r1 = batch_emu.plan(b, 1)
    assert r1["nsmall"] == 2 and r1["nwide"] == 4 and r1["nrounds"] == 1 + 6


def test_batch_tile_width_does_not_change_results():
    b = Batch()
    e, w = workloads.erdos_renyi_graph(14, 0.5, 21)
    b.add_qaoa(14, e, w, 2, [0.3, 1.1], [0.7, 0.2])
    terms = workloads.random_pauli_hamiltonian(14, 8, 4, 22)
    b.add_vqe(14, 1, terms, list(np.random.default_rng(1).uniform(-1, 1, 14)))
    got = {}
    for tq in (7, 11, 12, 13):
        if tq < 9:
            with pytest.raises(Exception):
                b.set_tile(tq)
            continue
        b.set_tile(tq)
        got[tq] = batch_emu.evaluate(b)[0]
    tol = 1e-5 * (1 + sum(w) + sum(abs(c) for c, _ in terms))
    assert np.allclose(got[12], got[13], rtol=0, atol=tol) and np.allclose(got[11], got[12], rtol=0, atol=tol)


def test_batch_wide_edge_tilings_on_cpu():
    """Smallest tile (2^9) with 17 and 18 qubits: tiling B with 1 and 0 low
    bits (every tile bit a high qubit), the XCD-swizzled tile order on 512
    tiles, QAOA energies taken in the last pass of either tiling."""
    b = Batch()
    qc = []
    for k, (n, p) in enumerate([(18, 1), (17, 2)]):
        e, w = workloads.erdos_renyi_graph(n, 0.3, 500 + k)
        rng = np.random.default_rng(70 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    b.set_tile(9, exact=True)
    en, am = batch_emu.evaluate(b)
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
        assert abs(en[k] - want) <= 1e-5 * (1 + sum(w)), (n, en[k], want)
        pr = np.abs(amps) ** 2
        assert pr[am[k]] >= pr.max() * (1 - 1e-4)


def test_batch_widens_tiles_that_would_cut_lines():
    """With base tile 10, a 20-qubit state's second tiling would read runs of
    one amplitude: it gets 12-qubit tiles (runs of 16 amplitudes, whole
    128-byte lines) while a 15-qubit state keeps 10; every pass then launches
    once per tile width.  Energies equal the oracle's, as with uniform tiles."""
    b = Batch()
    qc = []
    for k, (n, p) in enumerate([(15, 2), (20, 2)]):
        e, w = workloads.erdos_renyi_graph(n, 0.4, 600 + k)
        rng = np.random.default_rng(90 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    b.set_tile(10)
    assert list(batch_emu.data(b)["insts"]["tq"]) == [10, 12]
    r = batch_emu.plan(b, -1)
    assert list(r["launch_pass"]) == [0, 0, 1, 1, 2, 2] and list(r["launch_tq"]) == [12, 10] * 3
    assert list(r["pass_grid"]) == [1 << 8, 1 << 5] * 3
    assert r["ntiles"] == (1 << 8) + (1 << 5)
    en, am = batch_emu.evaluate(b)
    b.set_tile(10, exact=True)
    assert list(batch_emu.data(b)["insts"]["tq"]) == [10, 10]
    assert list(batch_emu.plan(b, -1)["launch_tq"]) == [10] * 3
    en_exact, _ = batch_emu.evaluate(b)
    for k, (n, e, w, p, g, be) in enumerate(qc):
        want, amps = orc.qaoa_cost(n, e, w, g, be, want_state=True)
        tol = 1e-5 * (1 + sum(w))
        assert abs(en[k] - want) <= tol and abs(en_exact[k] - want) <= tol, (n, en[k], en_exact[k], want)
        pr = np.abs(amps) ** 2
        assert pr[am[k]] >= pr.max() * (1 - 1e-4)


def test_batch_widened_vqe_state_on_cpu():
    """A VQE state past the config's 16 qubits: with base tile 10 an 18-qubit
    state gets 11-qubit tiles (runs of 16 amplitudes in the second tiling)
    beside a 14-qubit state on 10-qubit tiles; its Pauli energy is folded over
    the wider tiles.  Energies equal the oracle's, as with uniform tiles."""
    b = Batch()
    cases = []
    for k, n in enumerate((14, 18)):
        terms = workloads.random_pauli_hamiltonian(n, 8, 4, 700 + k)
        prm = list(np.random.default_rng(70 + k).uniform(-1, 1, 2 * n))
        cases.append((n, terms, prm))
        b.add_vqe(n, 2, terms, prm)
    b.set_tile(10)
    assert list(batch_emu.data(b)["insts"]["tq"]) == [10, 11]
    en, _ = batch_emu.evaluate(b)
    b.set_tile(10, exact=True)
    en_exact, _ = batch_emu.evaluate(b)
    for k, (n, terms, prm) in enumerate(cases):
        want = orc.energy(n, workloads.ry_cz_ansatz(n, 2, prm).to_array(), terms)
        tol = 1e-5 * (1 + sum(abs(c) for c, _ in terms))
        assert abs(en[k] - want) <= tol and abs(en_exact[k] - want) <= tol, (n, en[k], en_exact[k], want)


def test_batch_vqe_qubit_range():
    b = Batch()
    with pytest.raises(Exception):
        b.add_vqe(25, 1, workloads.random_pauli_hamiltonian(4, 2, 2, 1))
    with pytest.raises(Exception):
        b.add_vqe(1, 1, [(1.0, [(0, 3)])])
    b.add_vqe(24, 1, workloads.random_pauli_hamiltonian(24, 2, 2, 1))


def test_wide_rotation_factoring_modes():
    """Wide states rotate with the larger of cos / sin factored out of each
    2x2 (4 forms: Rx or Ry, cos or sin factored) and one uniform scale per
    pass; angles on both sides of pi/4 exercise all four, over several
    passes."""
    b = Batch()
    e, w = workloads.erdos_renyi_graph(14, 0.5, 77)
    g, be = [0.4, 1.1, 2.9], [0.3, 1.4, 2.6]  # beta: cos-dominant, sin-dominant, cos-dominant (negative)
    b.add_qaoa(14, e, w, 3, g, be)
    terms = workloads.random_pauli_hamiltonian(14, 10, 3, 78)
    prm = list(np.random.default_rng(79).uniform(-math.pi, math.pi, 14 * 2))  # half angles up to pi/2
    b.add_vqe(14, 2, terms, prm)
    en, am = batch_emu.evaluate(b, grid=2)
    want, amps = orc.qaoa_cost(14, e, w, g, be, want_state=True)
    assert abs(en[0] - want) <= 1e-5 * (1 + sum(w)), (en[0], want)
    pr = np.abs(amps) ** 2
    assert pr[am[0]] >= pr.max() * (1 - 1e-4)
    want_v = orc.energy(14, workloads.ry_cz_ansatz(14, 2, prm).to_array(), terms)
    assert abs(en[1] - want_v) <= 1e-5 * (1 + sum(abs(c) for c, _ in terms)), (en[1], want_v)


def test_shared_prefix_rounds_match_shifted_circuits():
    """Rounds 1.. of an optimizer iteration at fixed angles: the circuits
    shifted at gamma_l / beta_l start at pass l from the instance's running
    prefix state (advanced by one pass per round), and each must give the
    energy of the full circuit with that one angle shifted by +-pi/2."""
    b = Batch()
    qc = []
    for k, (n, p) in enumerate([(14, 3), (15, 2), (14, 1), (9, 2)]):
        e, w = workloads.erdos_renyi_graph(n, 0.5, 500 + k)
        rng = np.random.default_rng(90 + k)
        qc.append((n, e, w, p, list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))))
        b.add_qaoa(*qc[-1])
    b.add_vqe(14, 2, workloads.random_pauli_hamiltonian(14, 6, 3, 91))
    r0 = batch_emu.plan(b, 0)
    rounds = list(range(1, r0["nrounds"]))
    energy = batch_emu.evaluate_rounds(b, rounds)
    seen_shared = 0
    for rnd in rounds:
        r = batch_emu.plan(b, rnd)
        items = list(r["small"]) + list(r["wide"])
        first = dict(zip([(int(it["inst"]), int(it["variant"])) for it in r["wide"]], r["wfirst"][:r["nwide"]]))
        for it in items:
            if it["out"] < 0:
                continue
            k, j, var = int(it["inst"]), int(it["param"]), int(it["variant"])
            n, e, w, p, g, be = qc[k]
            g2, b2 = list(g), list(be)
            shift = math.pi / 2 if var == 1 else -math.pi / 2
            if j < p:
                g2[j] += shift
            else:
                b2[j - p] += shift
            want = orc.qaoa_cost(n, e, w, g2, b2)
            assert abs(energy[it["out"]] - want) <= 1e-5 * (1 + sum(w)), (rnd, k, j, var, energy[it["out"]], want)
            seen_shared += first.get((k, var), 0) > 0
    assert seen_shared > 0  # some shifted circuits did start from a shared prefix


def test_vqe_round0_chain_matches_shifted_circuits():
    """Round 0 with VQE states above 13 qubits: the base circuit runs as a
    chain of single-pass items keeping every pass's output, and the circuits
    shifted at layer l start at pass l from it; every energy must equal the
    oracle's at that one angle shifted by +-pi/2 (base: no shift)."""
    b = Batch()
    vc = []
    for k, (n, L) in enumerate([(14, 3), (15, 2), (12, 2)]):
        terms = workloads.random_pauli_hamiltonian(n, 8, 3, 700 + k)
        prm = list(np.random.default_rng(710 + k).uniform(-math.pi, math.pi, n * L))
        vc.append((n, L, terms, prm))
        b.add_vqe(n, L, terms, prm)
    e, w = workloads.erdos_renyi_graph(14, 0.5, 720)
    b.add_qaoa(14, e, w, 2, [0.3, 0.9], [1.2, 0.4])
    energy = batch_emu.evaluate_rounds(b, [0])
    r = batch_emu.plan(b, 0)
    firsts = {}
    for it, f in zip(r["wide"], r["wfirst"][:r["nwide"]]):
        if it["out"] >= 0:  # not the chain's single-pass items
            firsts[(int(it["inst"]), int(it["variant"]), int(it["param"]))] = int(f)
    checked = 0
    for it in list(r["small"]) + list(r["wide"]):
        k = int(it["inst"])
        if it["out"] < 0 or k >= len(vc):
            continue
        n, L, terms, prm = vc[k]
        p2 = list(prm)
        if it["variant"]:
            p2[int(it["param"])] += math.pi / 2 if it["variant"] == 1 else -math.pi / 2
        want = orc.energy(n, workloads.ry_cz_ansatz(n, L, p2).to_array(), terms)
        got = energy[it["out"]]
        assert abs(got - want) <= 1e-5 * (1 + sum(abs(c) for c, _ in terms)), (k, int(it["param"]), got, want)
        checked += 1
    assert checked == sum(1 + 2 * n * L for n, L, _, _ in vc)
    assert firsts[(0, 0, -1)] == 3 and firsts[(0, 1, 14 * 2)] == 2  # base from the chain; layer-2 shift at pass 2
