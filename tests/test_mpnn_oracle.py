"""Pins the warm-start network's FP64 oracle (oracle/mpnn_oracle.py): its
hand-written backward pass against central finite differences on sampled
parameters of every tensor, the aggregation operator's symmetry (the backward
reuses it as its own transpose), pooling and loss masking by depth."""
import numpy as np
import pytest

from oracle import mpnn_oracle as M
from ruqu_amd import workloads


def small_batch(seed=0, sizes=(5, 7, 4), ps=(1, 3, 8)):
    graphs = []
    rng = np.random.default_rng(seed)
    for k, n in enumerate(sizes):
        e, _ = workloads.erdos_renyi_graph(n, 0.5, 40 + 10 * seed + k)
        if not e:
            e = [(0, 1)]
        w = list(rng.uniform(0, 1, len(e)))
        graphs.append((n, e, w))
    return M.Batch(graphs, list(ps))


def test_layout_size():
    assert M.NPARAMS == M.KIN * M.H + M.H + 3 * (2 * M.H * M.H + M.H) + M.HIN * M.H + M.H + M.H * M.NOUT + M.NOUT


def test_aggregation_is_symmetric():
    b = small_batch()
    rng = np.random.default_rng(1)
    u, v = rng.normal(size=(b.nnodes, 3)), rng.normal(size=(b.nnodes, 3))
    assert np.isclose(np.sum(b.aggregate(u) * v), np.sum(u * b.aggregate(v)))


def test_loss_masks_unused_depths():
    b = small_batch()
    flat = M.init_params(3)
    O = M.forward(flat, b)
    T = np.zeros((b.ngraphs, 16))
    base, _ = M.loss_grad_out(O, T, b.p)
    T2 = T.copy()
    T2[0, 1:8] = 5.0  # graph 0 has p = 1: gammas 1..7 unused
    T2[0, 9:16] = 5.0
    assert M.loss_grad_out(O, T2, b.p)[0] == base


def test_backward_matches_finite_differences():
    b = small_batch()
    rng = np.random.default_rng(7)
    flat = M.init_params(5).astype(np.float64)
    flat += rng.normal(0, 0.01, flat.shape)  # break exact zeros of the biases
    T = rng.uniform(0, 1, (b.ngraphs, 16))
    loss, g = M.backward(flat, b, T)
    eps = 1e-6
    for name, (o, s) in M.OFFS.items():
        size = int(np.prod(s))
        for i in rng.choice(size, 6, replace=False):
            fp, fm = flat.copy(), flat.copy()
            fp[o + i] += eps
            fm[o + i] -= eps
            lp = M.loss_grad_out(M.forward(fp, b), T, b.p)[0]
            lm = M.loss_grad_out(M.forward(fm, b), T, b.p)[0]
            fd = (lp - lm) / (2 * eps)
            assert abs(fd - g[o + i]) <= 1e-6 + 1e-4 * abs(fd), (name, i, fd, g[o + i])


def test_bf16_rounding():
    # bf16 spacing on [1, 2) is 2^-7.  1 + 2^-8 is a tie (to even: 1.0); 1 + 2^-7 + 2^-8
    # is a tie whose lower neighbour is odd (to 1 + 2^-6); 1 + 2^-9 rounds down, 1 + 3*2^-9 up.
    x = np.array([1.0, 1.00390625, 1.01171875, 1.001953125, 1.005859375, -3.14159265, 0.0], dtype=np.float32)
    r = M.bf16(x)
    assert list(r[:5]) == [1.0, 1.0, 1.015625, 1.0, 1.0078125]
    assert r[5] == -3.140625 and r[6] == 0.0
