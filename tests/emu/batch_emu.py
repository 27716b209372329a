"""Runs the batched kernel's source on the CPU (tests/emu/libbatch_emu.so)
over the records and launch plan that ruqu_amd/csrc/batch.cpp hands the
device for rq_batch_evaluate, exported by rq_batch_plan_export /
rq_batch_data_export (host-only calls, no GPU).  Test infrastructure only."""
import ctypes as C
import os

import numpy as np

from ruqu_amd import _native

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libbatch_emu.so")

INST = np.dtype([("kind", "<i4"), ("n", "<i4"), ("p", "<i4"), ("nparams", "<i4"), ("nterms", "<i4"),
                 ("term_off", "<i4"), ("param_off", "<i4"), ("tq", "<i4"), ("ctab_off", "<i8"), ("wsum", "<f8")])
ITEM = np.dtype([("inst", "<i4"), ("variant", "<i4"), ("param", "<i4"), ("out", "<i4")])
EDGE = np.dtype([("i", "<u4"), ("j", "<u4"), ("w", "<f8")])
TERM = np.dtype([("x", "<u8"), ("zy", "<u8"), ("ny", "<u4"), ("pad", "<u4"), ("c", "<f8")])


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


PART = np.dtype([("e", "<f8"), ("pm", "<f8"), ("idx", "<u4"), ("pad", "<u4")])
I32P, I64P = C.POINTER(C.c_int32), C.POINTER(C.c_int64)


def _emu():
    from tests.emu import build_emu
    srcs = [build_emu.SRC, os.path.join(HERE, "batch_emu.cpp"), os.path.join(HERE, "hip_cpu_shim.h")]
    if not os.path.exists(LIB) or os.path.getmtime(LIB) < max(os.path.getmtime(f) for f in srcs):
        build_emu.build()
    lib = C.CDLL(LIB)
    vp, i = C.c_void_p, C.c_int
    lib.emu_cut_tables.argtypes = [vp, vp, vp, vp, i]
    lib.emu_batch_eval.argtypes = [vp, vp, i] + [vp] * 7 + [i]
    lib.emu_wide_pass.argtypes = [i] + [vp] * 9 + [i, i, vp] + [vp] * 4
    lib.emu_wide_energy.argtypes = [i] + [vp] * 9 + [vp]
    lib.emu_wide_fold.argtypes = [i] + [vp] * 5 + [vp, vp]
    return lib


def data(batch):
    """Instance, edge, term and parameter records of a ruqu_amd.batch.Batch."""
    lib = _native.lib()
    lib.rq_batch_data_export.argtypes = [C.c_void_p] * 5 + [C.POINTER(C.c_size_t)]
    cnt = (C.c_size_t * 4)()
    assert lib.rq_batch_data_export(batch._h, None, None, None, None, cnt) == 0
    insts, edges = np.zeros(cnt[0], dtype=INST), np.zeros(max(cnt[1], 1), dtype=EDGE)
    terms, params = np.zeros(max(cnt[2], 1), dtype=TERM), np.zeros(max(cnt[3], 1))
    assert lib.rq_batch_data_export(batch._h, _p(insts), _p(edges), _p(terms), _p(params), cnt) == 0
    return dict(insts=insts, edges=edges, terms=terms, params=params)


def plan(batch, rnd=-1):
    """The launch plan of one round (-1: plain evaluation) as batch.cpp builds it."""
    lib = _native.lib()
    lib.rq_batch_plan_export.argtypes = [C.c_void_p, C.c_int, I32P, I64P, C.c_void_p, C.c_void_p, I32P, I64P, I32P,
                                         I32P, I32P]
    info = np.zeros(8, dtype=np.int32)
    info64 = np.zeros(2, dtype=np.int64)
    ip = lambda a: a.ctypes.data_as(I32P)  # noqa: E731
    lp = lambda a: a.ctypes.data_as(I64P)  # noqa: E731
    assert lib.rq_batch_plan_export(batch._h, rnd, ip(info), lp(info64), None, None, None, None, None, None,
                                    None) == 0
    nsmall, nwide, ntiles, npasses = (int(x) for x in info[:4])
    nvtiles = int(info[7])
    r = dict(small=np.zeros(max(nsmall, 1), dtype=ITEM)[:nsmall], wide=np.zeros(max(nwide, 1), dtype=ITEM)[:nwide],
             wtile0=np.zeros(max(nwide, 1), dtype=np.int32), wstate=np.zeros(max(nwide, 1), dtype=np.int64),
             wmap=np.zeros(max(ntiles, 1), dtype=np.int32), pass_grid=np.zeros(max(npasses, 1), dtype=np.int32),
             vmap=np.zeros(max(nvtiles, 1), dtype=np.int32))
    assert lib.rq_batch_plan_export(batch._h, rnd, None, None, _p(r["small"]), _p(r["wide"]), ip(r["wtile0"]),
                                    lp(r["wstate"]), ip(r["wmap"]), ip(r["pass_grid"]), ip(r["vmap"])) == 0
    r.update(nsmall=nsmall, nwide=nwide, ntiles=ntiles, npasses=npasses, nvtiles=nvtiles, nrounds=int(info[4]),
             nouts=int(info[5]), tq=int(info[6]), state_amps=int(info64[0]), ctab_total=int(info64[1]))
    r["pass_grid"] = r["pass_grid"][:npasses]
    lib.rq_batch_plan_passes.argtypes = [C.c_void_p, C.c_int, I32P, I32P, I64P, I32P, I32P]
    total = int(r["pass_grid"].sum()) if npasses else 0
    r.update(wfirst=np.zeros(max(nwide, 1), dtype=np.int32), wlast=np.zeros(max(nwide, 1), dtype=np.int32),
             wsrc=np.zeros(max(nwide, 1), dtype=np.int64), ptile=np.zeros(max(total, 1), dtype=np.int32),
             pass_off=np.zeros(max(npasses, 1), dtype=np.int32))
    assert lib.rq_batch_plan_passes(batch._h, rnd, ip(r["wfirst"]), ip(r["wlast"]), lp(r["wsrc"]), ip(r["ptile"]),
                                    ip(r["pass_off"])) == 0
    lib.rq_batch_plan_launches.argtypes = [C.c_void_p, C.c_int, I32P, I32P]
    r.update(launch_pass=np.zeros(max(npasses, 1), dtype=np.int32), launch_tq=np.zeros(max(npasses, 1), dtype=np.int32))
    assert lib.rq_batch_plan_launches(batch._h, rnd, ip(r["launch_pass"]), ip(r["launch_tq"])) == 0
    return r


def _cut_tables(lib, insts, ea, total):
    ctab = np.zeros(max(total, 1), dtype=np.float32)
    # the launch shape of batch.cpp: BATCH_CUT_CHUNK = 4096 entries per workgroup
    chunks = np.array([(k, c) for k in range(len(insts)) if int(insts[k]["ctab_off"]) >= 0
                       for c in range(-(-(1 << int(insts[k]["n"])) // 4096))],
                      dtype=np.dtype([("inst", "<i4"), ("chunk", "<u4")]))
    if len(chunks):
        lib.emu_cut_tables(_p(ctab), _p(insts), _p(ea), _p(chunks), len(chunks))
    return ctab


def _run_round(lib, d, r, ctab, states, energy, amax, grid):
    """One round's launches (batch.cpp's enqueue_round) on host threads."""
    insts, ea, ta, pa = d["insts"], d["edges"], d["terms"], d["params"]
    done = np.zeros(len(insts), dtype=np.int32)
    if r["nsmall"]:
        lib.emu_batch_eval(_p(insts), _p(r["small"]), r["nsmall"], _p(done), _p(ea), _p(ta), _p(pa), _p(ctab),
                           _p(energy), _p(amax), min(grid, r["nsmall"]))
    if r["nwide"]:
        parts = np.zeros(r["ntiles"], dtype=PART)
        w = (_p(insts), _p(r["wide"]), _p(r["wmap"]), _p(r["wtile0"]), _p(r["wstate"]), _p(done))
        for li, g in enumerate(r["pass_grid"]):  # one launch per pass and tile width
            if g:
                lib.emu_wide_pass(int(g), *w, _p(pa), _p(ctab), _p(states), int(r["launch_tq"][li]),
                                  int(r["launch_pass"][li]), _p(parts), _p(r["ptile"][int(r["pass_off"][li]):]),
                                  _p(r["wfirst"]), _p(r["wlast"]), _p(r["wsrc"]))
        if r["nvtiles"]:
            lib.emu_wide_energy(r["nvtiles"], *w[:5], _p(r["vmap"]), _p(done), _p(ta), _p(states), _p(parts))
        lib.emu_wide_fold(r["nwide"], _p(insts), _p(r["wide"]), _p(r["wtile0"]), _p(done), _p(parts), _p(energy),
                          _p(amax))


def evaluate(batch, grid=3):
    """Energies and argmax indices per instance, as rq_batch_evaluate returns
    them: the launches of batch.cpp's enqueue_round run block after block on
    host threads over the exported plan."""
    d, r = data(batch), plan(batch, -1)
    lib = _emu()
    ctab = _cut_tables(lib, d["insts"], d["edges"], r["ctab_total"])
    energy = np.zeros(r["nouts"], dtype=np.float64)
    amax = np.zeros(r["nouts"], dtype=np.uint32)
    states = np.zeros(max(r["state_amps"], 1), dtype=np.complex64)
    _run_round(lib, d, r, ctab, states, energy, amax, grid)
    en = np.zeros(len(d["insts"]))
    am = np.zeros(len(d["insts"]), dtype=np.uint64)
    for it in list(r["small"]) + list(r["wide"]):
        if it["out"] >= 0:
            en[it["inst"]] = energy[it["out"]]
            am[it["inst"]] = amax[it["out"]]
    return en, am


def evaluate_rounds(batch, rounds, grid=3):
    """Rounds of one optimizer iteration, in the given order, at fixed
    parameters (no updates between them) and over one state buffer, so that
    the shared prefix states carry from round to round.  Returns the energy
    of every output slot."""
    d = data(batch)
    lib = _emu()
    plans = [plan(batch, rnd) for rnd in rounds]
    ctab = _cut_tables(lib, d["insts"], d["edges"], plans[0]["ctab_total"])
    nouts = plans[0]["nouts"]
    energy = np.zeros(nouts, dtype=np.float64)
    amax = np.zeros(nouts, dtype=np.uint32)
    states = np.zeros(max(max(p["state_amps"] for p in plans), 1), dtype=np.complex64)
    for r in plans:
        _run_round(lib, d, r, ctab, states, energy, amax, grid)
    return energy
