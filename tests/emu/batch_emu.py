"""Runs the batched kernel's source on the CPU (tests/emu/libbatch_emu.so)
over the records that ruqu_amd/csrc/batch.cpp hands the device for
rq_batch_evaluate: the instance, edge, term and parameter arrays and the
sorted work list, exported by rq_batch_plan_export / rq_batch_data_export
(host-only calls, no GPU).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from ruqu_amd import _native

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbatch_emu.so")

INST = np.dtype([("kind", "<i4"), ("n", "<i4"), ("p", "<i4"), ("nparams", "<i4"), ("nterms", "<i4"),
                 ("term_off", "<i4"), ("param_off", "<i4"), ("pad", "<i4"), ("ctab_off", "<i8"), ("wsum", "<f8")])
ITEM = np.dtype([("inst", "<i4"), ("variant", "<i4"), ("param", "<i4"), ("out", "<i4")])
EDGE = np.dtype([("i", "<u4"), ("j", "<u4"), ("w", "<f8")])
TERM = np.dtype([("x", "<u8"), ("zy", "<u8"), ("ny", "<u4"), ("pad", "<u4"), ("c", "<f8")])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def _emu():
    from tests.emu import build_emu
    srcs = [build_emu.SRC, os.path.join(HERE, "batch_emu.cpp"), os.path.join(HERE, "hip_cpu_shim.h")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(f) for f in srcs):
        build_emu.build()
    lib = C.CDLL(LIB)
    lib.emu_cut_table.argtypes = [C.c_void_p, C.c_int, C.c_void_p, C.c_int]
    lib.emu_batch_eval.argtypes = [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p, C.c_void_p, C.c_int]
    return lib


def export(batch):
    """The device records of a ruqu_amd.batch.Batch, as numpy arrays."""
    lib = _native.lib()
    lib.rq_batch_plan_export.argtypes = [C.c_void_p] * 3 + [C.POINTER(C.c_int)] * 2 + [C.POINTER(C.c_uint64),
                                                                                      C.POINTER(C.c_int64)]
    lib.rq_batch_data_export.argtypes = [C.c_void_p] * 4 + [C.POINTER(C.c_size_t)] * 3
    ni, nb, sa, ct = C.c_int(), C.c_int(), C.c_uint64(), C.c_int64()
    assert lib.rq_batch_plan_export(batch._h, None, None, C.byref(ni), C.byref(nb), C.byref(sa), C.byref(ct)) == 0
    insts, items = np.zeros(ni.value, dtype=INST), np.zeros(nb.value, dtype=ITEM)
    assert lib.rq_batch_plan_export(batch._h, _p(insts), _p(items), None, None, None, None) == 0
    ne, nt, npm = C.c_size_t(), C.c_size_t(), C.c_size_t()
    assert lib.rq_batch_data_export(batch._h, None, None, None, C.byref(ne), C.byref(nt), C.byref(npm)) == 0
    edges, terms, params = np.zeros(ne.value, dtype=EDGE), np.zeros(nt.value, dtype=TERM), np.zeros(npm.value)
    assert lib.rq_batch_data_export(batch._h, _p(edges), _p(terms), _p(params), None, None, None) == 0
    return dict(insts=insts, items=items, edges=edges, terms=terms, params=params, slot_amps=sa.value,
                ctab_total=ct.value)


def evaluate(batch, grid=4):
    """Energies and argmax indices per instance, as rq_batch_evaluate returns
    them, with the kernel run block after block on host threads."""
    r = export(batch)
    lib = _emu()
    insts, items = r["insts"], r["items"]
    ea = r["edges"] if len(r["edges"]) else np.zeros(1, dtype=EDGE)
    ta = r["terms"] if len(r["terms"]) else np.zeros(1, dtype=TERM)
    ctab = np.zeros(max(r["ctab_total"], 1), dtype=np.float32)
    for k in range(len(insts)):
        off = int(insts[k]["ctab_off"])
        if off >= 0:
            lib.emu_cut_table(_p(ctab[off:]), int(insts[k]["n"]), _p(ea[int(insts[k]["term_off"]):]),
                              int(insts[k]["nterms"]))
    nouts = int(items["out"].max()) + 1
    slots = np.zeros(max(grid * r["slot_amps"], 1), dtype=np.complex64)
    energy = np.zeros(nouts, dtype=np.float64)
    amax = np.zeros(nouts, dtype=np.uint32)
    done = np.zeros(len(insts), dtype=np.int32)
    lib.emu_batch_eval(_p(insts), _p(items), len(items), _p(done), _p(ea), _p(ta), _p(r["params"]), _p(ctab),
                       _p(slots), r["slot_amps"], _p(energy), _p(amax), grid)
    en = np.zeros(len(insts))
    am = np.zeros(len(insts), dtype=np.uint64)
    for it in items:
        en[it["inst"]] = energy[it["out"]]
        am[it["inst"]] = amax[it["out"]]
    return en, am
