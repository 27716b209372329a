"""Matrix-core phases of the fused schedule (complex f32) on the device.

A phase's ops are composed on the host into 16x16 unitaries and applied with
v_mfma_f32_32x32x2_f32.  Bars: max-abs 1e-5 against the FP64 oracle (the
f32 bar of the fused schedule) and 2e-6 against the host execution of the
same packed program (rq_fusion_emulate: identical operands, FP64
accumulation), at thresholds that turn the phases off, on wherever allowed,
and at the default.
"""
import ctypes as C

import numpy as np
import pytest

from ruqu_amd import SimConfig, Simulator, autotune, workloads
from ruqu_amd._native import lib, ptr
from tests.helpers import max_abs, mixed_circuit, orc

pytestmark = pytest.mark.gpu


def flush(msg):
    print(msg, flush=True)


@pytest.fixture
def mf_threshold():
    old = lib().rq_tune_get(4)
    yield lambda v: lib().rq_tune_set(4, v)
    lib().rq_tune_set(4, old)


def emulate(n, arr):
    amps = np.zeros(1 << n, dtype=np.complex64)
    amps[0] = 1
    assert lib().rq_fusion_emulate(n, 1, ptr(arr), len(arr), lib().rq_tune_get(2), lib().rq_tune_get(3),
                                   ptr(amps)) == 0, lib().rq_last_error()
    return amps


def packed_info(n, arr):
    r, f, u = C.c_uint64(), C.c_uint64(), C.c_uint64()
    assert lib().rq_fusion_packed_info(n, 1, ptr(arr), len(arr), lib().rq_tune_get(2), lib().rq_tune_get(3),
                                       C.byref(r), C.byref(f), C.byref(u)) == 0
    return r.value, f.value, u.value


def circuit(kind, n):
    return workloads.random_circuit(n, 8, 11 + n) if kind == "random" else mixed_circuit(n, 5, 300 + n)


@pytest.mark.parametrize("thr", [-1, 0, 768])
@pytest.mark.parametrize("kind,n", [("random", 12), ("random", 14), ("random", 18), ("mixed", 13), ("mixed", 16)])
def test_matrix_core_phases(mf_threshold, thr, kind, n):
    mf_threshold(thr)
    c = circuit(kind, n)
    arr = c.to_array()
    recs, mf, unitaries = packed_info(n, arr)
    flush("matrix-core phases: %s n=%d threshold %d -> %d phases (%d unitaries), %d records" %
          (kind, n, thr, mf, unitaries, recs))
    if thr == 0:
        assert mf > 0
    res = Simulator.run_with_config(c, SimConfig(precision="c32", fusion=True))
    got = res.state.state_vector()
    res.state.close()
    want, _, _ = orc.run_circuit(n, arr)
    e_orc = max_abs(got, want)
    e_emu = max_abs(got, emulate(n, arr))
    flush("  max-abs vs oracle %.3e, vs host program %.3e" % (e_orc, e_emu))
    assert e_orc < 1e-5
    assert e_emu < 2e-6


def test_autotune_picks_a_threshold(mf_threshold, tmp_path, monkeypatch):
    monkeypatch.setenv("RUQU_AMD_TUNE_CACHE", str(tmp_path / "tune.json"))
    m, c = lib().rq_tune_get(2), lib().rq_tune_get(3)
    try:
        res = autotune.tune_fused(14, "c32", candidates=[(12, 5)], mf_candidates=(-1, 768), use_cache=False)
        flush("tune_fused: %s" % res)
        assert res["mf"] in (-1, 768) and lib().rq_tune_get(4) == res["mf"]
        lib().rq_tune_set(4, 0)
        again = autotune.tune_fused(14, "c32")  # from the cache
        assert again["mf"] == res["mf"] and lib().rq_tune_get(4) == res["mf"]
    finally:
        lib().rq_tune_set(2, m)
        lib().rq_tune_set(3, c)


@pytest.mark.parametrize("m", [10, 11])
def test_matrix_core_phases_small_tiles(mf_threshold, m):
    """Tiles of 10 and 11 qubits: one or two waves per workgroup, so fewer
    (or no) wave bits can select among the composed unitaries."""
    old_m = lib().rq_tune_get(2)
    try:
        assert lib().rq_tune_set(2, m) == 0
        mf_threshold(0)
        for kind, n in (("random", 14), ("mixed", 13)):
            c = circuit(kind, n)
            arr = c.to_array()
            recs, mf, unitaries = packed_info(n, arr)
            flush("m=%d %s n=%d: %d matrix-core phases (%d unitaries)" % (m, kind, n, mf, unitaries))
            assert mf > 0
            res = Simulator.run_with_config(c, SimConfig(precision="c32", fusion=True))
            got = res.state.state_vector()
            res.state.close()
            want, _, _ = orc.run_circuit(n, arr)
            assert max_abs(got, want) < 1e-5
            assert max_abs(got, emulate(n, arr)) < 2e-6
    finally:
        lib().rq_tune_set(2, old_m)
