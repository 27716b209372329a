"""ctypes binding of libruqu_gpu.so (the C ABI declared in include/ruqu_gpu.h).

The library is the product: there is no fallback.  If it is missing, loading
raises ImportError naming the build step; if no GPU is present, every call
that needs the device returns RQ_ERR_DEVICE, which the Python layer raises as
DeviceError.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libruqu_gpu.so")

# rq_gate (88 bytes): kind, q0, q1, pad, theta, u[2][2] as (re, im) pairs.
GATE_DTYPE = np.dtype(
    [("kind", "<i4"), ("q0", "<u4"), ("q1", "<u4"), ("_pad", "<u4"), ("theta", "<f8"), ("u", "<f8", (8,))]
)
assert GATE_DTYPE.itemsize == 88

MEAS_DTYPE = np.dtype([("qubit", "<u4"), ("result", "<i4"), ("probability", "<f8")])
PAULI_DTYPE = np.dtype([("qubit", "<u4"), ("op", "<i4")])


class Metrics(C.Structure):
    _fields_ = [
        ("num_qubits", C.c_uint32),
        ("_pad", C.c_uint32),
        ("gate_count", C.c_uint64),
        ("execution_time_ns", C.c_uint64),
        ("peak_memory_bytes", C.c_uint64),
        ("gates_per_second", C.c_double),
        ("gates_fused", C.c_uint64),
        ("passes", C.c_uint64),
        ("device_time_ms", C.c_double),
        ("hbm_bytes", C.c_double),
    ]


class Noise(C.Structure):  # rq_noise (types.rs:249-253)
    _fields_ = [("depolarizing_rate", C.c_double), ("bit_flip_rate", C.c_double), ("phase_flip_rate", C.c_double)]


class Measurement(C.Structure):
    _fields_ = [("qubit", C.c_uint32), ("result", C.c_int32), ("probability", C.c_double)]


_vp = C.c_void_p
_u32 = C.c_uint32
_u64 = C.c_uint64
_sz = C.c_size_t
_i = C.c_int
_d = C.c_double
_pd = C.POINTER(C.c_double)

# name -> (restype, argtypes)
SIGNATURES = {
    "rq_last_error": (C.c_char_p, []),
    "rq_max_qubits": (_u32, [_i]),
    "rq_device_synchronize": (_i, []),
    "rq_state_create": (_i, [_u32, _i, _i, C.POINTER(_vp)]),
    "rq_state_from_amplitudes": (_i, [_vp, _sz, _u32, _i, C.POINTER(_vp)]),
    "rq_state_destroy": (None, [_vp]),
    "rq_state_num_qubits": (_u32, [_vp]),
    "rq_state_precision": (_i, [_vp]),
    "rq_estimate_memory": (_sz, [_u32, _i]),
    "rq_state_read": (_i, [_vp, _vp]),
    "rq_state_write": (_i, [_vp, _vp]),
    "rq_state_set_zero": (_i, [_vp]),
    "rq_state_copy": (_i, [_vp, _vp]),
    "rq_state_from_f64": (_i, [_vp, _vp]),
    "rq_state_to_f64": (_i, [_vp, _vp]),
    "rq_precision_error_bound": (_d, [_vp]),
    "rq_apply_gate": (_i, [_vp, _vp, _d, C.POINTER(Measurement), C.POINTER(_i)]),
    "rq_apply_1q": (_i, [_vp, _u32, _vp]),
    "rq_apply_2q": (_i, [_vp, _u32, _u32, _vp]),
    "rq_run_circuit": (_i, [_vp, _vp, _sz, _vp, _sz, _vp, _sz, C.POINTER(_sz), _i, C.POINTER(Metrics)]),
    "rq_apply_circuit": (_i, [_vp, _vp, _sz, _vp, _sz, _vp, _sz, C.POINTER(_sz), _i, C.POINTER(Metrics)]),
    "rq_run_circuit_noisy": (_i, [_vp, _vp, _sz, C.POINTER(Noise), _vp, _sz, C.POINTER(_sz), _vp, _sz,
                                 C.POINTER(_sz), _i, C.POINTER(Metrics)]),
    "rq_circuit_depth": (_u32, [_u32, _vp, _sz]),
    "rq_state_checkpoint_bytes": (_sz, [_vp]),
    "rq_state_checkpoint": (_i, [_vp, _vp]),
    "rq_state_from_checkpoint": (_i, [_vp, _sz, _i, C.POINTER(_vp)]),
    "rq_measure": (_i, [_vp, _u32, _d, C.POINTER(Measurement)]),
    "rq_measure_all": (_i, [_vp, _vp, _vp]),
    "rq_reset_qubit": (_i, [_vp, _u32, _d]),
    "rq_measurement_record": (_sz, [_vp, _vp, _sz]),
    "rq_sample_shots": (_i, [_vp, _vp, _sz, _vp, C.POINTER(_u64)]),
    "rq_stdrng_seed": (None, [_vp, _u64]),
    "rq_stdrng_next_u64": (_u64, [_vp]),
    "rq_stdrng_next_u32": (_u32, [_vp]),
    "rq_stdrng_next_f64": (None, [_vp, _vp, _sz]),
    "rq_stdrng_shot_draws": (None, [_u64, _sz, _sz, _vp]),
    "rq_chacha_block": (None, [_vp, _i, _vp]),
    "rq_probabilities": (_i, [_vp, _vp]),
    "rq_probability_of_qubit": (_i, [_vp, _u32, _pd]),
    "rq_expectation": (_i, [_vp, _vp, _sz, _pd]),
    "rq_expectation_hamiltonian": (_i, [_vp, _vp, _vp, _vp, _sz, _pd]),
    "rq_expectation_hamiltonian_noalloc": (_i, [_vp, _vp, _vp, _vp, _sz, _pd]),
    "rq_expectation_z_all": (_i, [_vp, _vp]),
    "rq_normalize": (_i, [_vp]),
    "rq_fidelity": (_i, [_vp, _vp, _pd]),
    "rq_norm_sq": (_i, [_vp, _pd]),
    "rq_fusion_configure": (_i, [_i, _i]),
    "rq_fusion_plan_info": (_i, [_u32, _vp, _sz, _i, _i, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "rq_fusion_emulate": (_i, [_u32, _i, _vp, _sz, _i, _i, _vp]),
    "rq_fusion_packed_info": (_i, [_u32, _i, _vp, _sz, _i, _i, C.POINTER(_u64), C.POINTER(_u64), C.POINTER(_u64)]),
    "rq_fusion_packed_kinds": (_i, [_u32, _i, _vp, _sz, _i, _i, _vp]),
    "rq_fusion_packed_records": (_i, [_u32, _i, _vp, _sz, _i, _i, _vp, _vp, _sz, _vp]),
    "rq_tune_set": (_i, [_i, _i]),
    "rq_tune_get": (_i, [_i]),
    "rq_time_1q": (_i, [_vp, _u32, _vp, _i, _pd]),
    "rq_time_copy": (_i, [_vp, _i, _pd]),
}

_lib = None


def lib():
    """Load libruqu_gpu.so once; raise ImportError when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            "libruqu_gpu.so not found at %s: build it with `python ruqu_amd/build.py` "
            "(or __graft_entry__.build())" % LIB_PATH
        )
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def register(name, restype, argtypes):
    """Declare an additional entry point (used by the algorithm modules)."""
    SIGNATURES[name] = (restype, argtypes)
    if _lib is not None:
        f = getattr(_lib, name)
        f.restype = restype
        f.argtypes = argtypes


def ptr(a):
    """Data pointer of a numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(C.c_void_p)
