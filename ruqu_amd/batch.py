"""Batched QAOA MaxCut / VQE instances on one GPU (rq_batch_* in the C ABI).

Each instance keeps the reference optimizer's semantics (run_qaoa,
qaoa.rs:217-308; run_vqe, vqe.rs:148-216, here with the Ry-CZ ansatz); the
circuits run in complex-f32 with FP64 reductions, all instances of a round
in one persistent kernel launch, one optimizer iteration per hipGraph replay.
"""
import ctypes as C

import numpy as np

from . import _native
from ._native import PAULI_DTYPE, lib, ptr
from .qaoa import RqGraph
from .state import check

_native.register("rq_batch_create", C.c_int, [C.POINTER(C.c_void_p)])
_native.register("rq_batch_destroy", None, [C.c_void_p])
_native.register("rq_batch_add_qaoa", C.c_int, [C.c_void_p, C.POINTER(RqGraph), C.c_uint32, C.c_void_p, C.c_void_p,
                                                C.POINTER(C.c_uint32)])
_native.register("rq_batch_add_vqe", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p,
                                               C.c_void_p, C.c_size_t, C.c_void_p, C.POINTER(C.c_uint32)])
_native.register("rq_batch_evaluate", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p])
_native.register("rq_batch_optimize", C.c_int, [C.c_void_p, C.c_uint32, C.c_double, C.c_double, C.c_double, C.c_int,
                                                C.POINTER(C.c_double)])
_native.register("rq_batch_result", C.c_int, [C.c_void_p, C.c_uint32, C.POINTER(C.c_double), C.POINTER(C.c_uint64),
                                              C.POINTER(C.c_uint32), C.POINTER(C.c_int), C.c_void_p, C.c_void_p,
                                              C.c_void_p])
_native.register("rq_batch_set_params", C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p])
_native.register("rq_batch_set_all_params", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t])
_native.register("rq_batch_size", C.c_uint32, [C.c_void_p])
_native.register("rq_batch_last_ms", C.c_double, [C.c_void_p])
_native.register("rq_batch_set_tile", C.c_int, [C.c_void_p, C.c_int])
_native.register("rq_batch_set_tile_exact", C.c_int, [C.c_void_p, C.c_int])
_native.register("rq_batch_maxcut", C.c_int, [C.c_void_p, C.c_void_p])
_native.register("rq_batch_results", C.c_int, [C.c_void_p] + [C.c_void_p] * 6)
_native.register("rq_batch_get_tile", C.c_int, [C.c_void_p])


class Batch:
    def __init__(self):
        h = C.c_void_p()
        check(lib().rq_batch_create(C.byref(h)))
        self._h = h
        self._kinds = []
        self._nparams = []
        self._n = []
        self._layers = []

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rq_batch_destroy(h)
            except Exception:
                pass
            self._h = None

    def __len__(self):
        return len(self._kinds)

    def add_qaoa(self, num_nodes, edges, weights, p, init_gammas=None, init_betas=None):
        ei = np.array([e[0] for e in edges], dtype=np.uint32)
        ej = np.array([e[1] for e in edges], dtype=np.uint32)
        w = np.array(weights, dtype=np.float64)
        g = RqGraph(num_nodes, len(ei), ptr(ei), ptr(ej), ptr(w))
        ig = None if init_gammas is None else np.ascontiguousarray(init_gammas, dtype=np.float64)
        ib = None if init_betas is None else np.ascontiguousarray(init_betas, dtype=np.float64)
        i = C.c_uint32()
        check(lib().rq_batch_add_qaoa(self._h, C.byref(g), p, ptr(ig), ptr(ib), C.byref(i)))
        self._kinds.append("qaoa")
        self._nparams.append(2 * p)
        self._n.append(num_nodes)
        self._layers.append(p)
        return i.value

    def add_vqe(self, n, layers, terms, init_params=None):
        """terms: [(coeff, [(qubit, op), ...]), ...] with op 0=I 1=X 2=Y 3=Z."""
        coeffs = np.array([c for c, _ in terms], dtype=np.float64)
        offs = np.zeros(len(terms) + 1, dtype=np.uint64)
        ops = []
        for k, (_, o) in enumerate(terms):
            offs[k + 1] = offs[k] + len(o)
            ops += [(q, op) for q, op in o]
        opa = np.array(ops, dtype=PAULI_DTYPE) if ops else np.zeros(0, dtype=PAULI_DTYPE)
        ip = None if init_params is None else np.ascontiguousarray(init_params, dtype=np.float64)
        i = C.c_uint32()
        check(lib().rq_batch_add_vqe(self._h, n, layers, ptr(coeffs), ptr(offs), ptr(opa), len(terms), ptr(ip),
                                     C.byref(i)))
        self._kinds.append("vqe")
        self._nparams.append(n * layers)
        self._n.append(n)
        self._layers.append(layers)
        return i.value

    def evaluate(self):
        e = np.zeros(len(self), dtype=np.float64)
        a = np.zeros(len(self), dtype=np.uint64)
        check(lib().rq_batch_evaluate(self._h, ptr(e), ptr(a)))
        return e, a

    def optimize(self, max_iterations, qaoa_lr=0.1, vqe_lr=0.1, vqe_threshold=1e-6, use_graph=True):
        ms = C.c_double()
        check(lib().rq_batch_optimize(self._h, max_iterations, qaoa_lr, vqe_lr, vqe_threshold, 1 if use_graph else 0,
                                      C.byref(ms)))
        self._max_iter = max_iterations
        return ms.value

    def result(self, i):
        best = C.c_double()
        bidx = C.c_uint64()
        iters = C.c_uint32()
        conv = C.c_int()
        params = np.zeros(self._nparams[i])
        bparams = np.zeros(self._nparams[i])
        hist = np.zeros(max(getattr(self, "_max_iter", 1), 1))
        check(lib().rq_batch_result(self._h, i, C.byref(best), C.byref(bidx), C.byref(iters), C.byref(conv),
                                    ptr(params), ptr(bparams), ptr(hist)))
        out = {"best": best.value, "iterations": iters.value, "converged": bool(conv.value), "params": params,
               "energy_history": hist[: iters.value].copy()}
        out.update(best_params=bparams)
        if self._kinds[i] == "qaoa":
            p = self._nparams[i] // 2
            out.update(best_index=int(bidx.value), optimal_gammas=params[:p], optimal_betas=params[p:],
                       best_gammas=bparams[:p], best_betas=bparams[p:],
                       best_bitstring=[bool((bidx.value >> v) & 1) for v in range(self._n[i])])
        return out

    def set_params(self, i, params):
        a = np.ascontiguousarray(params, dtype=np.float64)
        check(lib().rq_batch_set_params(self._h, i, ptr(a)))

    def set_all_params(self, params):
        """Every instance's parameters (a list in instance order) in one upload."""
        a = np.ascontiguousarray(np.concatenate([np.asarray(p, dtype=np.float64).ravel() for p in params])
                                 if len(params) else np.zeros(0), dtype=np.float64)
        check(lib().rq_batch_set_all_params(self._h, ptr(a) if a.size else None, a.size))

    def maxcut(self):
        """Exact MaxCut value of every QAOA instance (device brute force)."""
        out = np.zeros(len(self), dtype=np.float64)
        check(lib().rq_batch_maxcut(self._h, ptr(out)))
        return out

    def results(self):
        """Bulk results: dict of arrays (best, best_index, iterations,
        converged) and per-instance lists params / best_params."""
        n = len(self)
        best, bidx = np.zeros(n), np.zeros(n, dtype=np.uint64)
        iters, conv = np.zeros(n, dtype=np.uint32), np.zeros(n, dtype=np.int32)
        tot = int(sum(self._nparams))
        prm, bprm = np.zeros(max(tot, 1)), np.zeros(max(tot, 1))
        check(lib().rq_batch_results(self._h, ptr(best), ptr(bidx), ptr(iters), ptr(conv), ptr(prm), ptr(bprm)))
        offs = np.concatenate([[0], np.cumsum(self._nparams)]).astype(np.int64)
        return {"best": best, "best_index": bidx, "iterations": iters, "converged": conv.astype(bool),
                "params": [prm[offs[i]:offs[i + 1]] for i in range(n)],
                "best_params": [bprm[offs[i]:offs[i + 1]] for i in range(n)]}

    def set_tile(self, tq, exact=False):
        """LDS tile (2^tq amplitudes per workgroup) of the states above 13
        qubits; unless exact, a state whose second tiling would read runs
        shorter than a 128-byte line gets a wider tile (up to 13 qubits)."""
        f = lib().rq_batch_set_tile_exact if exact else lib().rq_batch_set_tile
        check(f(self._h, int(tq)))

    def tile(self):
        return int(lib().rq_batch_get_tile(self._h))

    def last_ms(self):
        return float(lib().rq_batch_last_ms(self._h))
