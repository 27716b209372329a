"""The C ABI library loads and exports every function include/ruqu_gpu.h
declares; host-only entry points behave without a GPU (no device work here)."""
import ctypes as C
import os
import re

import numpy as np

from ruqu_amd._native import LIB_PATH, lib, ptr
from ruqu_amd.gates import QuantumCircuit

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include", "ruqu_gpu.h")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rq_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported():
    names = declared()
    assert len(names) > 40
    L = C.CDLL(LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_host_only_entry_points():
    L = lib()
    assert L.rq_estimate_memory(20, 0) == 16_777_216
    c = QuantumCircuit(3).h(0).h(1).h(2).cnot(0, 1).cnot(1, 2)
    arr = c.to_array()
    assert L.rq_circuit_depth(3, ptr(arr), len(arr)) == 3
    assert L.rq_tune_set(99, 1) == 7  # RQ_ERR_ARGUMENT
    assert b"unknown" in L.rq_last_error()


def test_gate_record_layout():
    c = QuantumCircuit(2).rzz(0, 1, 0.25)
    a = c.to_array()
    assert a.itemsize == 88 and a["kind"][0] == 15 and a["q1"][0] == 1 and a["theta"][0] == 0.25
