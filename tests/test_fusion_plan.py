"""The fused schedule's planner, checked on the CPU: the plan is executed on
host memory with the device kernel's data mapping (rq_fusion_emulate) and
compared with the oracle for many circuits, tile sizes and low-qubit counts.
No GPU call is made here."""
import ctypes as C

import numpy as np
import pytest

from ruqu_amd import workloads
from ruqu_amd._native import lib, ptr
from tests.helpers import max_abs, mixed_circuit, orc


def emulate(n, arr, m, c, prec="c64"):
    amps = np.zeros(1 << n, dtype=np.complex128 if prec == "c64" else np.complex64)
    amps[0] = 1
    rc = lib().rq_fusion_emulate(n, 0 if prec == "c64" else 1, ptr(arr), len(arr), m, c, ptr(amps))
    assert rc == 0, lib().rq_last_error()
    return amps


def plan_info(n, arr, m, c):
    s, p, o = C.c_uint64(), C.c_uint64(), C.c_uint64()
    assert lib().rq_fusion_plan_info(n, ptr(arr), len(arr), m, c, C.byref(s), C.byref(p), C.byref(o)) == 0
    return s.value, p.value, o.value


@pytest.mark.parametrize("n", [4, 5, 6, 8, 10, 13])
@pytest.mark.parametrize("m,c", [(4, 0), (6, 2), (8, 5), (12, 5), (10, 3)])
def test_fused_plan_matches_oracle(n, m, c):
    arr = mixed_circuit(n, 4, 1000 + n * 13 + m).to_array()
    want, _, _ = orc.run_circuit(n, arr)
    got = emulate(n, arr, m, c)
    assert max_abs(got, want) < 1e-12


@pytest.mark.parametrize("n", [8, 12, 14])
def test_fused_plan_random_circuit_workload(n):
    arr = workloads.random_circuit(n, 6, 3).to_array()
    want, _, _ = orc.run_circuit(n, arr)
    assert max_abs(emulate(n, arr, 12, 5), want) < 1e-12
    got32 = emulate(n, arr, 12, 5, "c32")
    assert max_abs(got32, want) < 1e-5


def test_fused_plan_reduces_passes():
    arr = workloads.random_circuit(30, 40, 2).to_array()
    stages, phases, ops = plan_info(30, arr, 12, 5)
    assert stages < len(arr) / 10
    assert ops <= len(arr)


def packed_info(n, arr, m, c, prec):
    """(op records, matrix-core phases, composed unitaries) of the packed program."""
    r, f, u = C.c_uint64(), C.c_uint64(), C.c_uint64()
    assert lib().rq_fusion_packed_info(n, 0 if prec == "c64" else 1, ptr(arr), len(arr), m, c, C.byref(r),
                                       C.byref(f), C.byref(u)) == 0
    return r.value, f.value, u.value


@pytest.fixture
def mf_threshold():
    """Sets the matrix-core threshold (tuning key 4) for one test and restores it."""
    old = lib().rq_tune_get(4)
    yield lambda v: lib().rq_tune_set(4, v)
    lib().rq_tune_set(4, old)


@pytest.mark.parametrize("thr", [-1, 0, 512])
@pytest.mark.parametrize("n,depth,seed", [(10, 6, 3), (12, 6, 4), (14, 4, 5)])
def test_packed_program_matrix_core_phases(mf_threshold, thr, n, depth, seed):
    """The packed program (folds, matrix-core phases with the 32x32x2 lane
    layout spelled out on the host) against the oracle."""
    mf_threshold(thr)
    arr = workloads.random_circuit(n, depth, seed).to_array()
    m = 12 if n >= 12 else 10
    recs, mf, _ = packed_info(n, arr, m, 5, "c32")
    if thr == 0:
        assert mf > 0
    if thr == -1:
        assert mf == 0
    want, _, _ = orc.run_circuit(n, arr)
    assert max_abs(emulate(n, arr, m, 5, "c32"), want) < 1e-5
    # complex-f64 states never use the matrix-core phases
    assert packed_info(n, arr, m, 5, "c64")[1] == 0


def test_matrix_core_phases_with_uniform_conditions(mf_threshold):
    """Mixed circuits put CNOT controls and CZ bits on tile-id and wave bits:
    matrix-core phases then select among several composed unitaries."""
    mf_threshold(0)
    phases = unitaries = 0
    for seed in range(6):
        arr = mixed_circuit(13, 4, 1077 + seed).to_array()
        _, f, u = packed_info(13, arr, 12, 5, "c32")
        phases, unitaries = phases + f, unitaries + u
        want, _, _ = orc.run_circuit(13, arr)
        assert max_abs(emulate(13, arr, 12, 5, "c32"), want) < 1e-5
    assert phases > 0 and unitaries > phases


def test_planner_seed_search_knob():
    """Tuning key 5 (planner seeds): range checks, and every seed count gives a
    plan that the emulator runs to the oracle's state, with no more stages than
    the plain greedy scan."""
    old = lib().rq_tune_get(5)
    try:
        assert lib().rq_tune_set(5, -1) != 0 and lib().rq_tune_set(5, 5000) != 0
        arr = workloads.random_circuit(16, 8, 7).to_array()
        want, _, _ = orc.run_circuit(16, arr)
        stages, states = {}, {}
        for seeds in (0, 16, 128):
            assert lib().rq_tune_set(5, seeds) == 0 and lib().rq_tune_get(5) == seeds
            stages[seeds] = plan_info(16, arr, 10, 5)[0]
            states[seeds] = emulate(16, arr, 10, 5)
            assert max_abs(states[seeds], want) < 1e-12
        assert stages[128] < stages[0]
        # the emulator ran each plan: different schedules round differently
        assert not np.array_equal(states[0], states[128])
    finally:
        lib().rq_tune_set(5, old)


def test_factored_records_replace_dense_2x2():
    """Every 2x2 the emitter meets is applied in factored form (F with a unit
    diagonal or anti-diagonal, 8 FMAs per pair) with its diagonal part in the
    phase's pending-diagonal table; the table absorbs register-bit CZ and
    diagonal ops.  The emulated program equals the oracle either way."""
    n = 14
    arr = workloads.random_circuit(n, 10, 21).to_array()
    want, _, _ = orc.run_circuit(n, arr)
    cnt = np.zeros(9, dtype=np.uint64)
    assert lib().rq_fusion_packed_kinds(n, 0, ptr(arr), len(arr), 10, 4, ptr(cnt)) == 0
    u1, u1f, tables = int(cnt[0]), int(cnt[7]), int(cnt[8])
    assert u1f > 0 and tables > 0 and u1 < u1f
    assert max_abs(emulate(n, arr, 10, 4), want) < 1e-12
    got32 = emulate(n, arr, 10, 4, "c32")
    assert max_abs(got32, want) < 1e-5


def test_packed_records_export():
    """rq_fusion_packed_records lists the same program rq_fusion_packed_kinds
    counts: the kinds agree, every VALU phase holds an even number of records
    (odd ones are padded), and a too-small buffer is refused."""
    n = 14
    arr = workloads.random_circuit(n, 10, 21).to_array()
    cnt = np.zeros(9, dtype=np.uint64)
    assert lib().rq_fusion_packed_kinds(n, 1, ptr(arr), len(arr), 10, 4, ptr(cnt)) == 0
    cap = 1 << 14
    words = np.zeros(cap, dtype=np.uint32)
    phase = np.zeros(cap, dtype=np.int32)
    count = C.c_size_t()
    assert lib().rq_fusion_packed_records(n, 1, ptr(arr), len(arr), 10, 4, ptr(words), ptr(phase), cap,
                                          C.byref(count)) == 0
    k = count.value
    assert k == int(cnt.sum())
    words, phase = words[:k], phase[:k]
    kind = words & 7
    table = (kind == 2) & (((words >> 14) & 3) == 0)
    got = np.bincount(np.where(table, 8, kind), minlength=9)
    assert np.array_equal(got, cnt.astype(np.int64))
    _, per_phase = np.unique(phase[phase >= 0], return_counts=True)
    assert np.all(per_phase % 2 == 0)
    small = C.c_size_t()
    assert lib().rq_fusion_packed_records(n, 1, ptr(arr), len(arr), 10, 4, ptr(words), ptr(phase), k - 1,
                                          C.byref(small)) != 0
    assert small.value == k


@pytest.mark.parametrize("prec,thr", [("c64", -1), ("c32", -1), ("c32", 0)])
def test_carried_factors_and_cnot_phases(mf_threshold, prec, thr):
    """Pending per-qubit factors cross phase and stage boundaries (applied in
    the qubit's next register phase, folded into matrix-core unitaries, or
    applied in its last register phase), and a controlled CNOT after a pending
    factor on its target carries the relative phase in its record.  Deep
    circuits on small tiles (many stages and phases) equal the oracle, and
    table records are rare."""
    mf_threshold(thr)
    n, m, c = 13, (10 if thr == 0 else 8), 3  # matrix-core phases need 6 thread bits
    arr = workloads.random_circuit(n, 30, 44).to_array()
    if thr == 0:
        assert packed_info(n, arr, m, c, prec)[1] > 0
    want, _, _ = orc.run_circuit(n, arr)
    tol = 1e-12 if prec == "c64" else 2e-5
    assert max_abs(emulate(n, arr, m, c, prec), want) < tol
    p = 0 if prec == "c64" else 1
    cnt = np.zeros(9, dtype=np.uint64)
    assert lib().rq_fusion_packed_kinds(n, p, ptr(arr), len(arr), m, c, ptr(cnt)) == 0
    _, phases, _ = plan_info(n, arr, m, c)
    assert int(cnt[8]) * 4 < phases  # tables: far fewer than phases
    cap = 1 << 16
    words = np.zeros(cap, dtype=np.uint32)
    phase = np.zeros(cap, dtype=np.int32)
    count = C.c_size_t()
    assert lib().rq_fusion_packed_records(n, p, ptr(arr), len(arr), m, c, ptr(words), ptr(phase), cap,
                                          C.byref(count)) == 0
    w = words[:count.value]
    cx_phase = ((w & 7) == 4) & (((w >> 5) & 3) == 1)
    assert int(cx_phase.sum()) > 0
    mixed = mixed_circuit(n, 8, 4242).to_array()
    want, _, _ = orc.run_circuit(n, mixed)
    assert max_abs(emulate(n, mixed, m, c, prec), want) < tol
