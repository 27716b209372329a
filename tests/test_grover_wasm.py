"""Oracle pinning for Grover (grover.rs) and the host-side checks of the
WASM-shaped API (crates/ruqu-wasm/src/lib.rs).  CPU only.

Grover's oracle restatement is pinned by the textbook closed form: after k
iterations with M distinct marked states out of N, the marked probability is
sin^2((2k + 1) theta), theta = asin(sqrt(M / N)); by the reference's own
tests (test_algorithms.rs:177-320, grover.rs:204-229); and by the listed-twice
rule (an index listed twice is negated twice by the reference's loop).
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc
from ruqu_amd import wasm_api


@pytest.mark.parametrize("n,targets,k", [(3, [5], 1), (4, [7], 3), (5, [0, 9], 2), (6, [1, 2, 3], 4), (7, [100], 8),
                                         (3, [0, 2, 4, 6], 1), (8, [17], 12)])
def test_grover_oracle_closed_form(n, targets, k):
    _, _, _, p, it = orc.run_grover(n, targets, k, [0.5] * n)
    theta = math.asin(math.sqrt(len(targets) / (1 << n)))
    assert it == k
    assert abs(p - math.sin((2 * k + 1) * theta) ** 2) < 1e-12


def test_grover_optimal_iterations_kats():  # grover.rs:204-229, test_algorithms.rs:273-306
    assert orc.grover_optimal_iterations(3, 1) == 2
    assert orc.grover_optimal_iterations(2, 2) >= 1
    assert orc.grover_optimal_iterations(1, 1) >= 1
    assert 10 <= orc.grover_optimal_iterations(8, 1) <= 15
    assert 7 <= orc.grover_optimal_iterations(8, 2) <= 11
    assert 1 <= orc.grover_optimal_iterations(2, 1) <= 2
    # round(pi/4 / theta - 1/2) by hand for N = 256, M = 1: theta = asin(1/16)
    assert orc.grover_optimal_iterations(8, 1) == round(math.pi / 4 / math.asin(1 / 16) - 0.5)


@pytest.mark.parametrize("n,targets,it,seed_draw,bar", [(4, [7], None, 0.3, 0.8), (3, [5], None, 0.3, 0.8),
                                                         (3, [0], None, 0.3, 0.8), (3, [1, 5], None, 0.7, 0.7),
                                                         (3, [0, 2, 4, 6], None, 0.2, 0.5)])
def test_grover_reference_kats(n, targets, it, seed_draw, bar):  # test_algorithms.rs:180-259
    amps, m, found, p, iters = orc.run_grover(n, targets, it, [seed_draw] * n)
    assert p > bar
    assert 0 <= m < (1 << n)
    assert found == (m in targets)
    assert abs(np.sum(np.abs(amps) ** 2) - 1.0) < 1e-12


def test_grover_explicit_iterations():  # test_algorithms.rs:262-271
    _, _, _, _, iters = orc.run_grover(3, [3], 2, [0.5] * 3)
    assert iters == 2


def test_grover_target_listed_twice_is_no_oracle():
    # The oracle negates index 3 twice: the state stays uniform, and the
    # success probability sums probabilities()[3] twice.
    amps, _, _, p, _ = orc.run_grover(3, [3, 3], 2, [0.5] * 3)
    assert abs(p - 2.0 / 8.0) < 1e-15


def test_grover_target_out_of_range():
    with pytest.raises(RuntimeError):
        orc.run_grover(3, [8], 1, [0.5] * 3)


def test_wasm_limits_and_argument_errors():  # lib.rs:81-92, :412-447
    assert wasm_api.max_qubits() == 25
    with pytest.raises(ValueError, match="Qubit limit exceeded: 26 requested, max 25 in WASM"):
        wasm_api.WasmQuantumCircuit(26)
    with pytest.raises(ValueError, match="Qubit limit exceeded"):
        wasm_api.grover_search(26, [1])
    with pytest.raises(ValueError, match="gammas length mismatch: expected 2 \\(p\\), got 1"):
        wasm_api.qaoa_maxcut(3, [0, 1], 2, [0.1], [0.2, 0.3])
    with pytest.raises(ValueError, match="betas length mismatch: expected 1 \\(p\\), got 2"):
        wasm_api.qaoa_maxcut(3, [0, 1], 1, [0.1], [0.2, 0.3])
    with pytest.raises(ValueError, match="even number of elements"):
        wasm_api.qaoa_maxcut(3, [0, 1, 2], 1, [0.1], [0.2])
    with pytest.raises(ValueError, match="Qubit limit exceeded"):
        wasm_api.qaoa_maxcut(30, [0, 1], 1, [0.1], [0.2])
