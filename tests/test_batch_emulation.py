"""The batched evaluation kernel's source run on the CPU (host threads, one
per work-item, real barriers; tests/emu) over the records that batch.cpp
builds for the device, against the oracle.  Covers the kernel's persistent
work loop, both state placements (<= 13 qubits in LDS, more streamed through
LDS tiles from an HBM slot), the cut tables, the Pauli-term encoding and the
host packing, without a GPU.  Tolerance as on the device: complex-f32 states
with FP64 reductions, |dE| <= 1e-5 * (1 + sum |weights or coefficients|)."""
import math

import numpy as np

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
    r = batch_emu.export(b)
    ins, items = r["insts"], r["items"]
    assert list(ins["n"]) == [16, 6, 12]
    assert list(ins["nparams"]) == [4, 2, 36]
    assert ins[0]["ctab_off"] == 0 and ins[1]["ctab_off"] == -1 and ins[2]["ctab_off"] == -1
    assert r["ctab_total"] == 1 << 16 and r["slot_amps"] == 1 << 16
    assert list(items["inst"]) == [0, 2, 1]  # largest state first
    assert np.isclose(ins[0]["wsum"], sum(w))
    assert len(r["params"]) == 4 + 2 + 36 and np.all(r["params"][:6] == 0.5) and np.all(r["params"][6:] == 0.1)
