"""The batched workload at its stated size: 4096 QAOA MaxCut instances
(n ~ U{10..22}, G(n, 0.5), U(0,1) weights, p ~ U{1..5}) and 512 VQE instances
(n ~ U{12..16}, Ry-CZ ansatz with 4 layers, 100 Pauli strings of weight <= 4
with N(0,1) coefficients) in one batch, evaluated at seeded random angles.
Per instance the semantics are evaluate_qaoa_cost (qaoa.rs:179-201) and
evaluate_energy (vqe.rs:122-131); a sample of the instances, spanning both
kernel paths (states in LDS up to 13 qubits, tiled HBM passes above) and the
largest QAOA states, is compared with the FP64 oracle at
|dE| <= 1e-5 * (1 + sum |weights or coefficients|), the complex-f32 bar
scaled by the observable's norm."""
import math

import numpy as np
import pytest

from ruqu_amd import config4, workloads
from ruqu_amd.batch import Batch
from tests.helpers import orc

pytestmark = pytest.mark.gpu


def test_config4_full_batch_against_oracle():
    qa = config4.qaoa_instances()
    vq = config4.vqe_instances()
    assert len(qa) == 4096 and len(vq) == 512
    assert all(len(h) == 100 for _, _, h in vq)
    rng = np.random.default_rng(44)
    b = Batch()
    angles = []
    for n, e, w, p in qa:
        g, bt = list(rng.uniform(0, math.pi, p)), list(rng.uniform(0, math.pi, p))
        angles.append((g, bt))
        b.add_qaoa(n, e, w, p, g, bt)
    params = []
    for n, layers, h in vq:
        prm = list(rng.uniform(-1, 1, n * layers))
        params.append(prm)
        b.add_vqe(n, layers, h, prm)
    print("config 4: %d instances" % len(b), flush=True)
    en, _ = b.evaluate()
    print("  evaluate %.1f ms" % b.last_ms(), flush=True)
    assert np.all(np.isfinite(en))
    sizes = np.array([n for n, _, _, _ in qa])
    pick = [int(i) for i in np.flatnonzero(sizes <= 13)[:4]] + [int(i) for i in np.flatnonzero((sizes > 13) &
                                                                                                 (sizes <= 18))[:4]]
    pick += [int(i) for i in np.flatnonzero(sizes == sizes.max())[:2]]
    for i in pick:
        n, e, w, p = qa[i]
        want = orc.qaoa_cost(n, e, w, angles[i][0], angles[i][1])
        tol = 1e-5 * (1 + sum(w))
        print("  qaoa #%d n=%d p=%d E=%.8f oracle %.8f diff %.2e" % (i, n, p, en[i], want, en[i] - want), flush=True)
        assert abs(en[i] - want) <= tol
    for k in (0, 1, 2, 511):
        n, layers, h = vq[k]
        want = orc.energy(n, workloads.ry_cz_ansatz(n, layers, params[k]).to_array(), h)
        got = en[len(qa) + k]
        tol = 1e-5 * (1 + sum(abs(c) for c, _ in h))
        print("  vqe #%d n=%d E=%.8f oracle %.8f diff %.2e" % (k, n, got, want, got - want), flush=True)
        assert abs(got - want) <= tol
