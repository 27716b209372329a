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
