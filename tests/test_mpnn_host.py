"""Host side of the warm-start network against its specification (the FP64
oracle): parameter layout, initial parameters and the graph packing (CSR,
node and graph features) must agree exactly, so device results can be
compared with the oracle on the same inputs."""
import numpy as np

from oracle import mpnn_oracle as M
from ruqu_amd import mpnn, workloads


def graphs(seed=0, sizes=(5, 9, 13, 4)):
    rng = np.random.default_rng(seed)
    out = []
    for k, n in enumerate(sizes):
        e, _ = workloads.erdos_renyi_graph(n, 0.5, 60 + k + 10 * seed)
        if not e:
            e = [(0, 1)]
        out.append((n, e, list(rng.uniform(0, 1, len(e)))))
    return out


def test_layout_matches_oracle():
    assert mpnn.NPARAMS == M.NPARAMS
    for name, o, s in mpnn.LAYOUT:
        assert M.OFFS[name] == (o, s)


def test_init_matches_oracle():
    assert np.array_equal(mpnn.init_params(11), M.init_params(11))


def test_packing_matches_oracle():
    gs, ps = graphs(), [1, 4, 8, 2]
    a, b = mpnn.GraphBatch(gs, ps), M.Batch(gs, ps)
    assert np.array_equal(a.node_off, b.node_off)
    assert np.array_equal(a.row_ptr, b.row_ptr)
    assert np.array_equal(a.col, b.src)
    assert np.array_equal(a.w, b.w.astype(np.float32))
    assert np.array_equal(a.x, b.x) and np.array_equal(a.gfeat, b.gfeat) and np.array_equal(a.p, b.p)


def test_targets_array():
    T = mpnn.targets_array([[0.1, 0.2], [0.3]], [[0.5, 0.6], [0.7]])
    assert T.shape == (2, 16) and T[0, 1] == np.float32(0.2) and T[0, 9] == np.float32(0.6) and T[1, 1] == 0
