"""What the trained warm-start network gains over the method without it, on
held-out instances (G(n,0.5), n ~ U{10..22}, U(0,1) weights, p ~ U{1..8};
streams disjoint from the training instances), with the committed
parameters (ruqu_amd/data/warmstart_params.npy):

* the plain batched optimizer (run_qaoa's rule from 0.5, 16 iterations) is
  the method without it; the warm start (network prediction, the default
  start and the per-depth constant judged by one batched evaluation each,
  plain fallback below 99 % of the plain energy) must keep every instance at
  >= 99 % of the plain energy and run >= 5x the instances/s;
* the network's own angles must beat the optimizer's start (0.5) on raw
  approximation ratio and on the fraction already within 1 % of plain, and
  beat the per-depth constant predictor (the mean angles of the training
  records for each p) on both."""
import numpy as np
import pytest

from ruqu_amd import warmstart as W
from ruqu_amd.mpnn import Network

pytestmark = pytest.mark.gpu


def test_warm_start_gain():
    insts = W.held_out(1024)
    W.run_plain(insts[:32], 1)  # kernels loaded before timing
    base = W.plain_baseline(insts, 16)
    net = Network.load(W.default_params_path())
    const = W.ConstantPredictor.load()
    ens = W.evaluate([net, W.DefaultInit(), const], insts, 16, log=print, plain=base)
    alone = W.evaluate(net, insts, 16, log=print, plain=base)
    start = W.evaluate(W.DefaultInit(), insts, 16, log=print, plain=base)
    const_alone = W.evaluate(const, insts, 16, log=print, plain=base)
    assert ens["min_final_over_plain"] >= 0.99 - 1e-12
    assert ens["mean_ratio_warm_final"] >= ens["mean_ratio_plain"]
    assert ens["speedup"] >= 5.0
    assert alone["mean_ratio_warm_raw"] >= start["mean_ratio_warm_raw"] + 0.01
    assert alone["within_1pct_energy"] >= start["within_1pct_energy"] + 0.1
    # the graph network against the per-depth mean of its training records
    assert alone["within_1pct_energy"] > const_alone["within_1pct_energy"]
    assert alone["mean_ratio_warm_raw"] > const_alone["mean_ratio_warm_raw"]
    # rescue updates: the same raw candidates, fewer fallbacks, still >= 99 % of plain everywhere
    res = W.evaluate([net, W.DefaultInit(), const], insts, 16, log=print, plain=base, rescue_iters=6)
    assert res["within_1pct_energy"] == ens["within_1pct_energy"]
    assert res["rescued"] > 0 and res["fallbacks"] == ens["fallbacks"] - res["rescued"]
    assert res["min_final_over_plain"] >= 0.99 - 1e-12
    assert res["mean_ratio_warm_final"] >= ens["mean_ratio_warm_final"] - 1e-3
