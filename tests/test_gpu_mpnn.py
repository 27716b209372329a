"""Device forward / backward / Adam of the warm-start network against the
FP64 oracle run with bfloat16 rounding at the same points (so only FP32
accumulation order separates them): outputs within 2e-3 (relative to the
output scale), gradients within 2 % in norm with per-tensor cosine >= 0.999.
Also checks the bf16 effect itself against the unrounded FP64 network and
that training on the device lowers the loss."""
import numpy as np
import pytest

from oracle import mpnn_oracle as M
from ruqu_amd import mpnn
from tests.test_mpnn_host import graphs

pytestmark = pytest.mark.gpu


def batch_and_targets(seed, sizes, ps):
    gs = graphs(seed, sizes)
    rng = np.random.default_rng(100 + seed)
    T = rng.uniform(0, 1.5, (len(gs), 16)).astype(np.float32)
    return gs, ps, T


@pytest.mark.parametrize("sizes,ps", [((5, 9, 13, 4), (1, 4, 8, 2)),
                                      (tuple(range(10, 23)) * 5, tuple((np.arange(65) % 8) + 1))])
def test_forward_matches_oracle(sizes, ps):
    gs, ps, _ = batch_and_targets(1, sizes, list(ps))
    params = mpnn.init_params(3)
    params += np.random.default_rng(4).normal(0, 0.02, params.shape).astype(np.float32)
    net = mpnn.Network(params)
    got = net.forward(mpnn.GraphBatch(gs, ps))
    ob = M.Batch(gs, ps)
    want = M.forward(params, ob, bf=True)[:, :16]
    scale = 1.0 + np.max(np.abs(want))
    print("max |dev - oracle(bf16)| = %.3e (scale %.2f)" % (np.max(np.abs(got - want)), scale), flush=True)
    assert np.max(np.abs(got - want)) <= 2e-3 * scale
    exact = M.forward(params, ob, bf=False)[:, :16]
    assert np.max(np.abs(got - exact)) <= 5e-2 * scale  # bf16 operands vs FP64


def test_gradient_matches_oracle():
    gs, ps, T = batch_and_targets(2, (6, 11, 17, 22, 9, 14), [1, 3, 8, 5, 2, 7])
    params = mpnn.init_params(5)
    params += np.random.default_rng(6).normal(0, 0.02, params.shape).astype(np.float32)
    net = mpnn.Network(params)
    loss, g = net.gradient(mpnn.GraphBatch(gs, ps), T)
    wl, wg = M.backward(params, M.Batch(gs, ps), T, bf=True)
    assert abs(loss - wl) <= 1e-3 * (1 + abs(wl))
    rel = np.linalg.norm(g - wg) / np.linalg.norm(wg)
    print("loss %.6f oracle %.6f; gradient relative error %.3e" % (loss, wl, rel), flush=True)
    assert rel <= 2e-2
    for name, o, s in mpnn.LAYOUT:
        a, b = g[o:o + int(np.prod(s))], wg[o:o + int(np.prod(s))]
        if np.linalg.norm(b) > 1e-8:
            cos = float(np.dot(a, b) / (np.linalg.norm(a) * np.linalg.norm(b)))
            assert cos >= 0.999, (name, cos)


def test_training_lowers_loss():
    gs, ps, T = batch_and_targets(3, tuple(range(10, 23)) * 2, [(k % 8) + 1 for k in range(26)])
    net = mpnn.Network(seed=7)
    b = mpnn.GraphBatch(gs, ps)
    first = net.train_step(b, T, 1e-3)
    for _ in range(60):
        last = net.train_step(b, T, 1e-3)
    print("loss %.5f -> %.5f" % (first, last), flush=True)
    assert last < 0.5 * first
    # parameters moved on the device and read back coherently
    assert not np.array_equal(net.params(), mpnn.init_params(7))
