"""Warm-start graph network for QAOA angles (rq_mpnn_* in the C ABI).

No reference counterpart: the reference's QAOA (qaoa.rs:217-308) always
starts from gamma = beta = 0.5.  This network maps a weighted graph and a
depth p (1..8) to (gamma_1..p, beta_1..p); it runs and trains on the device
(bfloat16 matrix-core products, FP32 accumulation, hand-written backward,
Adam).  Architecture, features and loss are specified by the FP64
restatement oracle/mpnn_oracle.py, against which the device results are
tested.

  node features  [degree/16, weighted degree/8]         (32 columns, zero padded)
  graph features [onehot(p), n/22, edge density]         (32 columns, zero padded)
  h0 = relu(x W_in + b_in); 3 x { h += relu([h | A h] W_k + b_k) }
  O = relu([mean_v h | graph features] W_h1 + b_h1) W_h2 + b_h2
  gammas = O[:, :8], betas = O[:, 8:16]
"""
import ctypes as C

import numpy as np

from . import _native
from ._native import lib, ptr
from .state import check

H, KIN, GF, NOUT, PMAX, LAYERS = 256, 32, 32, 64, 8, 3
HIN = H + GF


class RqMpnnBatch(C.Structure):
    _fields_ = [("ngraphs", C.c_int32), ("nnodes", C.c_int32), ("nedges", C.c_int64), ("node_off", C.c_void_p),
                ("row_ptr", C.c_void_p), ("col", C.c_void_p), ("w", C.c_void_p), ("x", C.c_void_p),
                ("gfeat", C.c_void_p), ("p", C.c_void_p)]


_native.register("rq_mpnn_num_params", C.c_size_t, [])
_native.register("rq_mpnn_create", C.c_int, [C.POINTER(C.c_void_p)])
_native.register("rq_mpnn_destroy", None, [C.c_void_p])
_native.register("rq_mpnn_set_params", C.c_int, [C.c_void_p, C.c_void_p])
_native.register("rq_mpnn_get_params", C.c_int, [C.c_void_p, C.c_void_p])
_native.register("rq_mpnn_forward", C.c_int, [C.c_void_p, C.POINTER(RqMpnnBatch), C.c_void_p])
_native.register("rq_mpnn_gradient", C.c_int, [C.c_void_p, C.POINTER(RqMpnnBatch), C.c_void_p,
                                               C.POINTER(C.c_float), C.c_void_p])
_native.register("rq_mpnn_train_step", C.c_int, [C.c_void_p, C.POINTER(RqMpnnBatch), C.c_void_p, C.c_float,
                                                 C.POINTER(C.c_float)])


def layout():
    """(name, offset, shape) of the flat FP32 parameters."""
    out, o = [], 0
    shapes = [("W_in", (KIN, H)), ("b_in", (H,))]
    for k in range(LAYERS):
        shapes += [("W_%d" % (k + 1), (2 * H, H)), ("b_%d" % (k + 1), (H,))]
    shapes += [("W_h1", (HIN, H)), ("b_h1", (H,)), ("W_h2", (H, NOUT)), ("b_h2", (NOUT,))]
    for name, s in shapes:
        out.append((name, o, s))
        o += int(np.prod(s))
    return out, o


LAYOUT, NPARAMS = layout()


def init_params(seed):
    """He-normal weights (numpy PCG64(seed)), zero biases, a small output
    layer and output biases 0.5, so a fresh network predicts angles near the
    plain optimizer's initial point."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(NPARAMS)
    for name, o, s in LAYOUT:
        if name.startswith("W"):
            flat[o:o + int(np.prod(s))] = rng.normal(0.0, np.sqrt(2.0 / s[0]), int(np.prod(s)))
        if name == "W_h2":
            flat[o:o + int(np.prod(s))] *= 0.1
        if name == "b_h2":
            flat[o:o + 2 * PMAX] = 0.5
    return flat.astype(np.float32)


class GraphBatch:
    """Graphs packed for the network.  graphs: [(n, edges, weights)], ps:
    depths.  Every undirected edge is stored in both directions; the edge
    list is CSR by destination with sources ascending."""

    def __init__(self, graphs, ps):
        self.ngraphs = len(graphs)
        self.p = np.ascontiguousarray(ps, dtype=np.int32)
        ns = np.array([g[0] for g in graphs], dtype=np.int64)
        self.node_off = np.zeros(self.ngraphs + 1, dtype=np.int32)
        self.node_off[1:] = np.cumsum(ns)
        self.nnodes = int(self.node_off[-1])
        src, dst, w = [], [], []
        x = np.zeros((self.nnodes, KIN), dtype=np.float32)
        gfeat = np.zeros((self.ngraphs, GF), dtype=np.float32)
        for gi, ((n, edges, weights), p) in enumerate(zip(graphs, ps)):
            base = int(self.node_off[gi])
            e = np.asarray(edges, dtype=np.int64).reshape(-1, 2)
            ww = np.asarray(weights, dtype=np.float64)
            src += [base + e[:, 0], base + e[:, 1]]
            dst += [base + e[:, 1], base + e[:, 0]]
            w += [ww, ww]
            deg = np.bincount(e.reshape(-1), minlength=n).astype(np.float64)
            wdeg = np.bincount(e[:, 0], weights=ww, minlength=n) + np.bincount(e[:, 1], weights=ww, minlength=n)
            x[base:base + n, 0] = deg / 16.0
            x[base:base + n, 1] = wdeg / 8.0
            gfeat[gi, p - 1] = 1.0
            gfeat[gi, PMAX] = n / 22.0
            gfeat[gi, PMAX + 1] = len(e) / (n * (n - 1) / 2.0)
        src = np.concatenate(src) if src else np.zeros(0, dtype=np.int64)
        dst = np.concatenate(dst) if dst else np.zeros(0, dtype=np.int64)
        w = np.concatenate(w) if w else np.zeros(0)
        order = np.lexsort((src, dst))
        self.col = np.ascontiguousarray(src[order], dtype=np.int32)
        self.w = np.ascontiguousarray(w[order], dtype=np.float32)
        self.row_ptr = np.zeros(self.nnodes + 1, dtype=np.int32)
        self.row_ptr[1:] = np.cumsum(np.bincount(dst, minlength=self.nnodes))
        self.x = x
        self.gfeat = gfeat

    def c_struct(self):
        return RqMpnnBatch(self.ngraphs, self.nnodes, len(self.col), ptr(self.node_off), ptr(self.row_ptr),
                           ptr(self.col), ptr(self.w), ptr(self.x), ptr(self.gfeat), ptr(self.p))


def targets_array(gammas_list, betas_list):
    """[G, 16] FP32 targets from per-graph angle lists (unused entries 0)."""
    T = np.zeros((len(gammas_list), 2 * PMAX), dtype=np.float32)
    for i, (g, b) in enumerate(zip(gammas_list, betas_list)):
        T[i, :len(g)] = g
        T[i, PMAX:PMAX + len(b)] = b
    return T


class Network:
    def __init__(self, params=None, seed=0):
        h = C.c_void_p()
        check(lib().rq_mpnn_create(C.byref(h)))
        self._h = h
        self.set_params(init_params(seed) if params is None else params)

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rq_mpnn_destroy(h)
            except Exception:
                pass
            self._h = None

    def set_params(self, params):
        a = np.ascontiguousarray(params, dtype=np.float32)
        assert a.size == NPARAMS == lib().rq_mpnn_num_params()
        check(lib().rq_mpnn_set_params(self._h, ptr(a)))

    def params(self):
        out = np.zeros(NPARAMS, dtype=np.float32)
        check(lib().rq_mpnn_get_params(self._h, ptr(out)))
        return out

    def forward(self, batch):
        """[G, 16] outputs: gammas in columns 0..7, betas in 8..15."""
        out = np.zeros((batch.ngraphs, 2 * PMAX), dtype=np.float32)
        bs = batch.c_struct()
        check(lib().rq_mpnn_forward(self._h, C.byref(bs), ptr(out)))
        return out

    def predict(self, graphs, ps):
        """Per graph: (gammas[:p], betas[:p]) as float64 lists."""
        out = self.forward(GraphBatch(graphs, ps))
        return [(out[i, :p].astype(np.float64), out[i, PMAX:PMAX + p].astype(np.float64)) for i, p in enumerate(ps)]

    def gradient(self, batch, targets):
        T = np.ascontiguousarray(targets, dtype=np.float32)
        g = np.zeros(NPARAMS, dtype=np.float32)
        loss = C.c_float()
        bs = batch.c_struct()
        check(lib().rq_mpnn_gradient(self._h, C.byref(bs), ptr(T), C.byref(loss), ptr(g)))
        return float(loss.value), g

    def train_step(self, batch, targets, lr):
        T = np.ascontiguousarray(targets, dtype=np.float32)
        loss = C.c_float()
        bs = batch.c_struct()
        check(lib().rq_mpnn_train_step(self._h, C.byref(bs), ptr(T), float(lr), C.byref(loss)))
        return float(loss.value)

    def save(self, path):
        np.save(path, self.params())

    @classmethod
    def load(cls, path):
        return cls(params=np.load(path))
