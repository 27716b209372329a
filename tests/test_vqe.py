"""VQE (vqe.rs) on the device against the oracle's run_vqe, plus the
reference's own VQE tests (vqe.rs:288-330, tests/test_algorithms.rs:325-460,
876-893).  Exact per-gate kernels give bit-identical f64 amplitudes; the
energy is an FP64 device reduction in a different summation order from the
reference's sequential sum, so energies agree to 1e-12 and every optimizer
trajectory is compared at that bar."""
import numpy as np
import pytest

from ruqu_amd import vqe, workloads
from tests.helpers import orc


def _terms(h):
    return [(c, [(q, int(op)) for q, op in p.ops]) for c, p in h.terms]


def test_num_parameters():  # vqe.rs:292-297
    assert vqe.num_parameters(2, 1) == 4
    assert vqe.num_parameters(4, 3) == 24
    assert vqe.num_parameters(1, 5) == 10


def test_build_ansatz_gate_count():  # vqe.rs:299-308
    c = vqe.build_ansatz(3, 2, [0.0] * vqe.num_parameters(3, 2))
    assert c.num_qubits() == 3
    assert len(c.gates()) == 16


def test_build_ansatz_wrong_param_count():  # vqe.rs:310-314 (panic -> ValueError)
    with pytest.raises(ValueError, match="expected 4 parameters"):
        vqe.build_ansatz(2, 1, [0.0] * 3)


def test_build_ansatz_matches_oracle():
    rng = np.random.default_rng(3)
    prm = rng.uniform(-3, 3, vqe.num_parameters(5, 3))
    a = vqe.build_ansatz(5, 3, prm).to_array()
    b = orc.build_ansatz(5, 3, prm)
    assert len(a) == len(b)
    for f in ("kind", "q0", "q1", "theta"):
        assert np.array_equal(a[f], b[f])


def test_hamiltonian_helpers():  # vqe.rs:316-328
    h = vqe.h2_hamiltonian()
    assert h.num_qubits == 2 and len(h.terms) == 5
    z = vqe.single_z_hamiltonian()
    assert z.num_qubits == 1 and len(z.terms) == 1 and z.terms[0][0] == -1.0


def _check_against_oracle(cfg):
    got = vqe.run_vqe(cfg)
    want = orc.run_vqe(cfg.num_qubits, cfg.ansatz_depth, _terms(cfg.hamiltonian), cfg.max_iterations,
                       cfg.convergence_threshold, cfg.learning_rate)
    assert got.num_iterations == want["num_iterations"]
    assert got.converged == want["converged"]
    assert np.max(np.abs(got.energy_history - want["energy_history"])) < 1e-12
    assert abs(got.optimal_energy - want["optimal_energy"]) < 1e-12
    assert np.max(np.abs(got.optimal_parameters - want["optimal_parameters"])) < 1e-10
    return got


@pytest.mark.gpu
@pytest.mark.parametrize("depth,iters,thr,lr", [(2, 50, 0.01, 0.1), (2, 100, 0.01, 0.05), (2, 20, 0.001, 0.1),
                                                (2, 30, 0.01, 0.1), (1, 10, 0.1, 0.1)])
def test_vqe_h2_matches_oracle(depth, iters, thr, lr):
    cfg = vqe.VqeConfig(vqe.h2_hamiltonian(), 2, depth, iters, thr, lr, seed=42)
    got = _check_against_oracle(cfg)
    assert np.isfinite(got.optimal_energy) and len(got.optimal_parameters) > 0
    if (depth, iters, thr, lr) == (2, 50, 0.01, 0.1):
        assert got.optimal_energy < -0.8  # test_vqe_h2_energy
    if (depth, iters, thr, lr) == (2, 20, 0.001, 0.1) and got.num_iterations >= 2:
        assert got.energy_history[-1] <= got.energy_history[0] + 1e-10  # test_vqe_energy_decreases


@pytest.mark.gpu
def test_vqe_simple_z():  # test_algorithms.rs:348-391
    from ruqu_amd.gates import Hamiltonian, PauliOp, PauliString
    h = Hamiltonian([(1.0, PauliString([(0, PauliOp.Z)]))], 1)
    got = _check_against_oracle(vqe.VqeConfig(h, 1, 1, 30, 0.01, 0.1, seed=42))
    assert -1.0 - 1e-10 <= got.optimal_energy <= 1.0 + 1e-10


@pytest.mark.gpu
def test_vqe_random_hamiltonian_matches_oracle():
    terms = workloads.random_pauli_hamiltonian(6, 20, 3, 11)
    from ruqu_amd.gates import Hamiltonian, PauliString
    h = Hamiltonian([(c, PauliString(ops)) for c, ops in terms], 6)
    _check_against_oracle(vqe.VqeConfig(h, 6, 2, 4, 1e-9, 0.05))


def test_vqe_evaluate_energy_requires_device_library():
    # The product path has no CPU fallback: without a GPU the state cannot be created.
    import torch  # noqa: F401  (only to learn whether a GPU is present)
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    from ruqu_amd.state import QuantumError
    with pytest.raises(QuantumError):
        vqe.evaluate_energy(vqe.VqeConfig(vqe.h2_hamiltonian(), 2, 1, 1, 0.1, 0.1), [0.1] * 4)
