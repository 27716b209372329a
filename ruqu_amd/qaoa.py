"""QAOA MaxCut: the API of crates/ruqu-algorithms/src/qaoa.rs on the device.

Graph, QaoaConfig, QaoaResult, build_qaoa_circuit, cut_value,
evaluate_qaoa_cost and run_qaoa keep the reference's names and meaning.
The work runs in libruqu_gpu.so: a one-pass cost-phase kernel per layer, the
Rx mixer (exact per-qubit kernels, or the fused tile schedule), FP64 energy
reductions and the plain optimizer loop in C++.

mode="exact": amplitudes bit-identical to the reference's f64 arithmetic;
mode="fused": fused mixer and one-pass energy (rounding-level differences).
"""
import ctypes as C

import numpy as np

from . import _native
from ._native import lib, ptr
from .gates import QuantumCircuit
from .state import QuantumState, _as_prec, check

_MODES = {"exact": 0, "fused": 1, 0: 0, 1: 1}


class RqGraph(C.Structure):
    _fields_ = [("num_nodes", C.c_uint32), ("num_edges", C.c_uint32), ("ei", C.c_void_p), ("ej", C.c_void_p),
                ("w", C.c_void_p)]


class RqQaoaResult(C.Structure):
    _fields_ = [("best_cut_value", C.c_double), ("best_index", C.c_uint64), ("iterations", C.c_uint32),
                ("converged", C.c_int32), ("evaluations", C.c_uint64), ("device_ms", C.c_double),
                ("wall_ms", C.c_double)]


_native.register("rq_qaoa_create", C.c_int, [C.POINTER(RqGraph), C.c_int, C.POINTER(C.c_void_p)])
_native.register("rq_qaoa_destroy", None, [C.c_void_p])
_native.register("rq_qaoa_state", C.c_void_p, [C.c_void_p])
_native.register("rq_qaoa_cost", C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_int,
                                           C.POINTER(C.c_double)])
_native.register("rq_qaoa_run", C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_double, C.c_int, C.c_void_p,
                                          C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                          C.POINTER(RqQaoaResult)])
_native.register("rq_argmax_probability", C.c_int, [C.c_void_p, C.POINTER(C.c_uint64), C.POINTER(C.c_double)])
_native.register("rq_expectation_zmasks", C.c_int, [C.c_void_p, C.c_void_p, C.c_size_t, C.c_void_p])


class Graph:
    """qaoa.rs:34-75."""

    def __init__(self, num_nodes, edges=None):
        self.num_nodes = int(num_nodes)
        self.edges = []
        for e in edges or []:
            self.add_edge(*e)

    def add_edge(self, i, j, weight=1.0):
        if not 0 <= i < self.num_nodes:
            raise ValueError("node index %d out of range" % i)
        if not 0 <= j < self.num_nodes:
            raise ValueError("node index %d out of range" % j)
        self.edges.append((int(i), int(j), float(weight)))

    @classmethod
    def unweighted(cls, num_nodes, edges):
        g = cls(num_nodes)
        g.edges = [(int(i), int(j), 1.0) for i, j in edges]
        return g

    def num_edges(self):
        return len(self.edges)

    def arrays(self):
        ei = np.array([e[0] for e in self.edges], dtype=np.uint32)
        ej = np.array([e[1] for e in self.edges], dtype=np.uint32)
        w = np.array([e[2] for e in self.edges], dtype=np.float64)
        return ei, ej, w


def triangle_graph():  # qaoa.rs:318-320
    return Graph.unweighted(3, [(0, 1), (1, 2), (0, 2)])


def ring4_graph():  # qaoa.rs:325-327
    return Graph.unweighted(4, [(0, 1), (1, 2), (2, 3), (3, 0)])


class QaoaConfig:
    """qaoa.rs:82-93."""

    def __init__(self, graph, p, max_iterations, learning_rate, seed=None):
        self.graph = graph
        self.p = int(p)
        self.max_iterations = int(max_iterations)
        self.learning_rate = float(learning_rate)
        self.seed = seed


class QaoaResult:
    """qaoa.rs:96-110 (+ evaluations and device time)."""

    def __init__(self, **kw):
        self.__dict__.update(kw)


def build_qaoa_circuit(graph, gammas, betas):  # qaoa.rs:125-153
    if len(gammas) != len(betas):
        raise ValueError("gammas and betas must have equal length")
    c = QuantumCircuit(graph.num_nodes)
    for q in range(graph.num_nodes):
        c.h(q)
    for layer in range(len(gammas)):
        for i, j, w in graph.edges:
            c.rzz(i, j, 2.0 * gammas[layer] * w)
        for q in range(graph.num_nodes):
            c.rx(q, 2.0 * betas[layer])
    return c


def cut_value(graph, bitstring):  # qaoa.rs:162-169
    return sum(w for i, j, w in graph.edges if bool(bitstring[i]) != bool(bitstring[j]))


def cut_values(graph, bits):
    """Cut value of every row of a [shots, n] 0/1 array (vectorised)."""
    bits = np.asarray(bits, dtype=np.uint8)
    out = np.zeros(bits.shape[0], dtype=np.float64)
    for i, j, w in graph.edges:
        out += w * (bits[:, i] != bits[:, j])
    return out


class QaoaInstance:
    """One graph with its device state (rq_qaoa)."""

    def __init__(self, graph, precision="c64"):
        self.graph = graph
        self.precision = precision
        self._ei, self._ej, self._w = graph.arrays()
        g = RqGraph(graph.num_nodes, len(self._ei), ptr(self._ei), ptr(self._ej), ptr(self._w))
        h = C.c_void_p()
        check(lib().rq_qaoa_create(C.byref(g), _as_prec(precision), C.byref(h)))
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value:
            try:
                lib().rq_qaoa_destroy(h)
            except Exception:
                pass
            self._h = None

    def cost(self, gammas, betas, mode="exact"):
        g = np.ascontiguousarray(gammas, dtype=np.float64)
        b = np.ascontiguousarray(betas, dtype=np.float64)
        if len(g) != len(b):
            raise ValueError("gammas and betas must have equal length")
        out = C.c_double()
        check(lib().rq_qaoa_cost(self._h, ptr(g), ptr(b), len(g), _MODES[mode], C.byref(out)))
        return out.value

    def state(self):
        """A non-owning QuantumState view of the instance's state."""
        st = QuantumState.__new__(QuantumState)
        st._h = C.c_void_p(lib().rq_qaoa_state(self._h))
        st._prec = _as_prec(self.precision)
        st._rng = np.random.default_rng()
        st._owner = self  # keep the instance alive; the handle is not destroyed by the view
        st.__class__ = _StateView
        return st

    def run(self, p, max_iterations, learning_rate, mode="exact", init_gammas=None, init_betas=None):
        g = np.zeros(p)
        b = np.zeros(p)
        hist = np.zeros(max(max_iterations, 1))
        ig = None if init_gammas is None else np.ascontiguousarray(init_gammas, dtype=np.float64)
        ib = None if init_betas is None else np.ascontiguousarray(init_betas, dtype=np.float64)
        res = RqQaoaResult()
        check(lib().rq_qaoa_run(self._h, p, max_iterations, learning_rate, _MODES[mode], ptr(ig), ptr(ib), ptr(g),
                                ptr(b), ptr(hist), C.byref(res)))
        n = self.graph.num_nodes
        return QaoaResult(best_cut_value=res.best_cut_value,
                          best_bitstring=[bool((res.best_index >> v) & 1) for v in range(n)],
                          best_index=int(res.best_index), optimal_gammas=g, optimal_betas=b,
                          energy_history=hist[: res.iterations].copy(), converged=bool(res.converged),
                          evaluations=int(res.evaluations), device_ms=res.device_ms, wall_ms=res.wall_ms)

    def sample_cuts(self, draws):
        """Sampled bitstrings of the current state (measure_all per shot with
        the given draws, shape [shots, n]) and their cut values."""
        bits, ties = self.state().sample_shots(draws)
        return bits, cut_values(self.graph, bits), ties


class _StateView(QuantumState):
    def __del__(self):
        self._h = None


def evaluate_qaoa_cost(graph, gammas, betas, seed=None, precision="c64", mode="exact"):  # qaoa.rs:179-201
    return QaoaInstance(graph, precision).cost(gammas, betas, mode)


def run_qaoa(config, precision="c64", mode="exact", init_gammas=None, init_betas=None):  # qaoa.rs:217-308
    inst = QaoaInstance(config.graph, precision)
    return inst.run(config.p, config.max_iterations, config.learning_rate, mode, init_gammas, init_betas)
