"""Learned warm start for batched QAOA MaxCut.

The plain method is the reference's optimizer, run_qaoa (qaoa.rs:217-308):
gamma = beta = 0.5, parameter-shift ascent with learning rate 0.1, the
result being the highest expected cut seen (best_cut_value).  Here it runs
for many instances at once on the batched engine (ruqu_amd.batch).

The warm start replaces the iterations by a prediction: the graph network
(ruqu_amd.mpnn) maps (graph, p) to angles, one batched evaluation gives the
expected cut there (optionally followed by a few plain iterations from
those angles), and the plain optimizer from the default start is the
fallback for every instance whose approximation ratio (expected cut /
exact MaxCut) misses the plain result's by more than 1 %.

Instances: n ~ U{10..22}, G(n, 0.5) with U(0,1) weights, p ~ U{1..8}, from
SplitMix64 streams of base seed 5 (training instances i >= 0, held-out
instances from index 2^29 on: disjoint streams).  Training records are the
plain optimizer's best angles on training instances.
"""
import time

import numpy as np

from . import workloads
from .batch import Batch
from .mpnn import PMAX, GraphBatch, Network, targets_array
from .workloads import SplitMix64, stream_seed

BASE = 5
HELD_OUT = 1 << 29
GRAPH_STREAM = 1 << 30


def default_params_path():
    """The shipped network: trained on the 200 000 plain-optimizer records
    (16 iterations each) of training instances 0..199 999 of this module's
    stream, 30 epochs, Adam lr 1e-3, 256 graphs per step."""
    import os
    return os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_params.npy")


def instances(count, start=0, base=BASE, nmin=10, nmax=22, pmax=PMAX):
    """[(n, edges, weights, p)] for instance indices start .. start+count-1."""
    out = []
    for i in range(start, start + count):
        rng = SplitMix64(stream_seed(base, i))
        n = nmin + rng.below(nmax - nmin + 1)
        p = 1 + rng.below(pmax)
        edges, w = workloads.erdos_renyi_graph(n, 0.5, stream_seed(base, GRAPH_STREAM + i))
        out.append((n, edges, w, p))
    return out


def held_out(count, base=BASE, **kw):
    return instances(count, start=HELD_OUT, base=base, **kw)


def _batch(insts, init=None, tile=None):
    b = Batch()
    for k, (n, e, w, p) in enumerate(insts):
        g0, b0 = (None, None) if init is None else init[k]
        b.add_qaoa(n, e, w, p, g0, b0)
    if tile:
        b.set_tile(tile)
    return b


def run_plain(insts, iterations, lr=0.1, init=None, tile=None):
    """Batched run_qaoa (from init when given).  Returns the bulk results
    plus wall_s (setup, optimization and read-back) and device_ms."""
    t0 = time.perf_counter()
    b = _batch(insts, init, tile)
    ms = b.optimize(iterations, qaoa_lr=lr, use_graph=True)
    r = b.results()
    r["wall_s"] = time.perf_counter() - t0
    r["device_ms"] = ms
    r["batch"] = b
    return r


def evaluate_at(insts, angles, tile=None):
    """Expected cut of every instance at the given (gammas, betas): one
    batched evaluation.  Returns (energies, wall_s, batch)."""
    t0 = time.perf_counter()
    b = _batch(insts, angles, tile)
    en, _ = b.evaluate()
    return en, time.perf_counter() - t0, b


def make_records(count, iterations, start=0, chunk=4096, lr=0.1, log=print):
    """Training records from the batched plain optimizer: per instance the
    graph, p, the angles of its best expected cut and that cut."""
    graphs, ps, gam, bet, best = [], [], [], [], []
    done = 0
    while done < count:
        k = min(chunk, count - done)
        insts = instances(k, start=start + done)
        r = run_plain(insts, iterations, lr)
        for (n, e, w, p), bp, bv in zip(insts, r["best_params"], r["best"]):
            graphs.append((n, e, w))
            ps.append(p)
            gam.append(bp[:p].copy())
            bet.append(bp[p:].copy())
            best.append(bv)
        done += k
        if log:
            log("records %d/%d (chunk device %.1f s, wall %.1f s)" % (done, count, r["device_ms"] / 1e3, r["wall_s"]))
    return {"graphs": graphs, "p": np.array(ps), "gammas": gam, "betas": bet, "best": np.array(best)}


def train(net, records, epochs, batch_graphs=256, lr=1e-3, seed=0, log=print):
    """Adam on the masked MSE between predicted and recorded angles.
    Mini-batches are packed once; their order is reshuffled every epoch."""
    idx = np.arange(len(records["p"]))
    rng = np.random.default_rng(seed)
    rng.shuffle(idx)
    batches = []
    for s in range(0, len(idx), batch_graphs):
        sel = idx[s:s + batch_graphs]
        gb = GraphBatch([records["graphs"][i] for i in sel], [int(records["p"][i]) for i in sel])
        T = targets_array([records["gammas"][i] for i in sel], [records["betas"][i] for i in sel])
        batches.append((gb, T))
    hist = []
    for ep in range(epochs):
        order = rng.permutation(len(batches))
        tot = 0.0
        for j in order:
            tot += net.train_step(batches[j][0], batches[j][1], lr)
        hist.append(tot / len(batches))
        if log:
            log("epoch %d loss %.5f" % (ep, hist[-1]))
    return hist


def predict(net, insts, chunk=2048):
    out = []
    for s in range(0, len(insts), chunk):
        part = insts[s:s + chunk]
        out += net.predict([(n, e, w) for n, e, w, _ in part], [p for *_, p in part])
    return out


def plain_baseline(insts, plain_iters, lr=0.1):
    """The plain method on the instances: batched run_qaoa from 0.5, its
    wall time, and the exact MaxCut of every instance."""
    rp = run_plain(insts, plain_iters, lr)
    rp["cmax"] = rp["batch"].maxcut()
    rp["ratio"] = rp["best"] / rp["cmax"]
    return rp


def evaluate(predictors, insts, plain_iters, refine_iters=0, lr=0.1, tol=0.01, log=print, plain=None,
             rescue_iters=0):
    """Plain batched QAOA vs the warm start on the same instances.

    predictors: one predictor (an object with predict(graphs, ps)) or a
    list; each candidate costs one batched evaluation and the simulator's
    expected cut picks the best per instance.  refine_iters > 0 then runs
    that many plain updates from the chosen angles (best cut seen).  An
    instance falls back to the plain optimizer from 0.5 when its warm
    expected cut is below (1 - tol) x the plain result's (within 1 % of the
    plain-optimized energy; approximation ratios share the exact MaxCut
    denominator); the fallback runs are timed as a separate batch.
    rescue_iters > 0 first gives the missing instances that many plain
    updates from their chosen angles (one batch of just those instances);
    only the ones still below the bar fall back.
    Instances/s = count / wall time of each pipeline."""
    if not isinstance(predictors, (list, tuple)):
        predictors = [predictors]
    N = len(insts)
    rp = plain if plain is not None else plain_baseline(insts, plain_iters, lr)
    cmax, ratio_plain = rp["cmax"], rp["ratio"]
    t_pred = t_eval = 0.0
    best_e = np.full(N, -np.inf)
    best_a = [None] * N
    chosen = np.zeros(N, dtype=np.int32)
    b = None  # one batch for every candidate: new angles are one upload
    for ci, pr in enumerate(predictors):
        t0 = time.perf_counter()
        cand = predict(pr, insts)
        t_pred += time.perf_counter() - t0
        t0 = time.perf_counter()
        if b is None:
            b = _batch(insts, cand)
        else:
            b.set_all_params([np.concatenate([np.asarray(g), np.asarray(bt)]) for g, bt in cand])
        e, _ = b.evaluate()
        t_eval += time.perf_counter() - t0
        for i in range(N):
            if e[i] > best_e[i]:
                best_e[i], best_a[i], chosen[i] = e[i], cand[i], ci
    del b  # frees the candidates' device states before the next batches
    if refine_iters > 0:
        rw = run_plain(insts, refine_iters + 1, lr, init=best_a)
        best_e = np.maximum(best_e, rw["best"])
        t_eval += rw["wall_s"]
    ratio_raw = best_e / cmax
    ok = best_e >= (1.0 - tol) * rp["best"]
    ok_raw = ok.copy()
    t_rescue, rescued = 0.0, 0
    miss = [i for i in range(N) if not ok[i]]
    if rescue_iters > 0 and miss:
        rr = run_plain([insts[i] for i in miss], rescue_iters + 1, lr, init=[best_a[i] for i in miss])
        t_rescue = rr["wall_s"]
        for j, i in enumerate(miss):
            if rr["best"][j] > best_e[i]:
                best_e[i] = rr["best"][j]
        ok = best_e >= (1.0 - tol) * rp["best"]
        rescued = int(np.sum(ok) - np.sum(ok_raw))
    ratio_warm = best_e / cmax
    fb = [i for i in range(N) if not ok[i]]
    t_fb = 0.0
    if fb:
        rf = run_plain([insts[i] for i in fb], plain_iters, lr)
        t_fb = rf["wall_s"]
    final = np.where(ok, ratio_warm, ratio_plain)
    res = {
        "instances": N,
        "candidates": len(predictors),
        "plain_iterations": plain_iters,
        "refine_iterations": refine_iters,
        "plain_wall_s": rp["wall_s"],
        "plain_device_ms": rp["device_ms"],
        "plain_inst_per_s": N / rp["wall_s"],
        "warm_predict_s": t_pred,
        "warm_eval_s": t_eval,
        "rescue_iterations": rescue_iters,
        "rescue_s": t_rescue,
        "rescued": rescued,
        "fallback_s": t_fb,
        "fallbacks": len(fb),
        "warm_inst_per_s": N / (t_pred + t_eval + t_rescue + t_fb),
        "mean_ratio_plain": float(np.mean(ratio_plain)),
        "mean_ratio_warm_raw": float(np.mean(ratio_raw)),
        "mean_ratio_warm_final": float(np.mean(final)),
        "within_1pct_energy": float(np.mean(ok_raw)),
        "within_1pt_ratio": float(np.mean(ratio_raw >= ratio_plain - 0.01)),
        "chosen_fraction": [float(np.mean(chosen == c)) for c in range(len(predictors))],
        "min_final_over_plain": float(np.min(final / ratio_plain)),
    }
    res["speedup"] = res["warm_inst_per_s"] / res["plain_inst_per_s"]
    if log:
        log("plain %.1f inst/s (ratio %.4f) | warm[%d] %.1f inst/s (final ratio %.4f, raw %.4f; %.1f%% within 1%% of "
            "plain energy, %d rescued by %d updates, %d fallbacks) -> %.2fx" % (
                res["plain_inst_per_s"], res["mean_ratio_plain"], len(predictors), res["warm_inst_per_s"],
                res["mean_ratio_warm_final"], res["mean_ratio_warm_raw"], 100 * res["within_1pct_energy"], rescued,
                rescue_iters, len(fb), res["speedup"]))
    return res


class DefaultInit:
    """The plain optimizer's own start (0.5 for every angle), evaluated
    once: the warm start with no learned input at all."""

    def predict(self, graphs, ps):
        return [(np.full(p, 0.5), np.full(p, 0.5)) for p in ps]


class ConstantPredictor:
    """Per-depth constant angles (ruqu_amd/data/warmstart_constant.json:
    the mean best angles of the training records for each p)."""

    def __init__(self, table):
        self.table = {int(p): (np.asarray(v["gammas"]), np.asarray(v["betas"])) for p, v in table.items()}

    @classmethod
    def load(cls, path=None):
        import json
        import os
        path = path or os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "warmstart_constant.json")
        with open(path) as f:
            return cls(json.load(f))

    def predict(self, graphs, ps):
        return [self.table[p] for p in ps]


def constant_table(records):
    """Mean recorded angles per depth, in the format of
    ruqu_amd/data/warmstart_constant.json (0.5 for a depth with no record)."""
    table = {}
    ps = np.asarray(records["p"])
    for p in range(1, PMAX + 1):
        sel = np.flatnonzero(ps == p)
        if len(sel):
            g = np.mean([records["gammas"][i] for i in sel], axis=0)
            b = np.mean([records["betas"][i] for i in sel], axis=0)
        else:
            g, b = np.full(p, 0.5), np.full(p, 0.5)
        table[str(p)] = {"gammas": [float(x) for x in g], "betas": [float(x) for x in b], "records": int(len(sel))}
    return table


def constant_predictor(records):
    """The method without the network: the mean recorded angles per depth
    (same for every graph), as an ablation of what the graph input adds."""
    return ConstantPredictor(constant_table(records))


def save_records(path, rec):
    ns = np.array([g[0] for g in rec["graphs"]], dtype=np.int32)
    ne = np.array([len(g[1]) for g in rec["graphs"]], dtype=np.int32)
    edges = np.concatenate([np.asarray(g[1], dtype=np.int16).reshape(-1, 2) for g in rec["graphs"]])
    w = np.concatenate([np.asarray(g[2], dtype=np.float64) for g in rec["graphs"]])
    G = np.zeros((len(ns), PMAX))
    B = np.zeros((len(ns), PMAX))
    for i, p in enumerate(rec["p"]):
        G[i, :p] = rec["gammas"][i]
        B[i, :p] = rec["betas"][i]
    np.savez_compressed(path, n=ns, ne=ne, edges=edges, w=w, p=rec["p"], gammas=G, betas=B, best=rec["best"])


def save_records_compact(path, rec, start):
    """Records of the consecutive training instances start .. start+count-1
    without their graphs: the graphs are a function of the index
    (instances()), so only p, the angles and the best cut are stored."""
    G = np.zeros((len(rec["p"]), PMAX))
    B = np.zeros((len(rec["p"]), PMAX))
    for i, p in enumerate(rec["p"]):
        G[i, :p] = rec["gammas"][i]
        B[i, :p] = rec["betas"][i]
    np.savez_compressed(path, start=np.int64(start), p=np.asarray(rec["p"], dtype=np.int8), gammas=G, betas=B,
                        best=np.asarray(rec["best"]))


def load_records_compact(paths):
    """Concatenate compact record files, regenerating each graph from its
    instance index; the regenerated depth must equal the stored one."""
    out = {"graphs": [], "p": [], "gammas": [], "betas": [], "best": []}
    for path in paths:
        with np.load(path) as z:
            arrs = {k: z[k] for k in z.files}
        p = arrs["p"].astype(np.int64)
        insts = instances(len(p), start=int(arrs["start"]))
        for i, (n, e, w, pi) in enumerate(insts):
            if pi != p[i]:
                raise ValueError("%s: record %d has p=%d, instance %d has p=%d" % (path, i, p[i], i, pi))
            out["graphs"].append((n, e, w))
            out["gammas"].append(arrs["gammas"][i, :pi])
            out["betas"].append(arrs["betas"][i, :pi])
        out["p"].append(p)
        out["best"].append(arrs["best"])
    out["p"] = np.concatenate(out["p"])
    out["best"] = np.concatenate(out["best"])
    return out


def subset(rec, count):
    """The first count records."""
    return {"graphs": rec["graphs"][:count], "p": rec["p"][:count], "gammas": rec["gammas"][:count],
            "betas": rec["betas"][:count], "best": rec["best"][:count]}


def load_records(path):
    with np.load(path) as z:  # each z[...] access decompresses: read every array once
        arrs = {k: z[k] for k in z.files}
    edges, w = arrs["edges"].astype(np.int64), arrs["w"]
    graphs, off = [], 0
    for n, k in zip(arrs["n"], arrs["ne"]):
        graphs.append((int(n), edges[off:off + k].tolist(), w[off:off + k].tolist()))
        off += k
    p = arrs["p"]
    return {"graphs": graphs, "p": p, "gammas": [arrs["gammas"][i, :p[i]] for i in range(len(p))],
            "betas": [arrs["betas"][i, :p[i]] for i in range(len(p))], "best": arrs["best"]}


def main(argv=None):
    """Records -> training -> held-out evaluation, all on the device."""
    import argparse
    import json
    import os
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--records", type=int, default=200000)
    ap.add_argument("--start", type=int, default=0, help="first training instance index of new records")
    ap.add_argument("--records-only", action="store_true",
                    help="make records (saved compact, without graphs) and stop")
    ap.add_argument("--record-file", default=None, help="reuse records saved by an earlier run")
    ap.add_argument("--compact-records", nargs="*", default=None,
                    help="reuse compact record files (graphs regenerated from the instance index)")
    ap.add_argument("--curve", default=None,
                    help="comma-separated record counts: train one network per prefix and evaluate each")
    ap.add_argument("--params", default=None, help="start from (or, with --epochs 0, evaluate) these parameters")
    ap.add_argument("--iters", type=int, default=16, help="plain optimizer iterations")
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval", type=int, default=4096)
    ap.add_argument("--out", default="warmstart_out")
    ap.add_argument("--seed", type=int, default=0)
    a = ap.parse_args(argv)
    os.makedirs(a.out, exist_ok=True)

    def log(msg):
        print(msg, flush=True)

    t0 = time.time()
    rec = None
    if a.compact_records:
        rec = load_records_compact(a.compact_records)
        log("loaded %d compact records (%.0f s)" % (len(rec["p"]), time.time() - t0))
    elif a.record_file:
        rec = load_records(a.record_file)
        log("loaded %d records" % len(rec["p"]))
    elif a.records > 0:
        rec = make_records(a.records, a.iters, start=a.start, log=log)
        if a.records_only:
            save_records_compact(os.path.join(a.out, "records_%d_%d.npz" % (a.start, a.records)), rec, a.start)
            log("records done in %.0f s" % (time.time() - t0))
            return
        save_records(os.path.join(a.out, "records.npz"), rec)
    if a.curve:
        return learning_curve(rec, [int(c) for c in a.curve.split(",")], a, log)
    t1 = time.time()
    net = Network.load(a.params) if a.params else Network(seed=a.seed)
    hist = train(net, rec, a.epochs, lr=a.lr, seed=a.seed, log=log) if (a.epochs and rec) else []
    net.save(os.path.join(a.out, "params.npy"))
    t2 = time.time()
    ev = held_out(a.eval)
    res = {"records": len(rec["p"]) if rec else 0, "plain_iterations": a.iters, "epochs": a.epochs,
           "train_loss": hist, "records_s": t1 - t0, "train_s": t2 - t1}
    base = plain_baseline(ev, a.iters)
    log("plain baseline: %d instances, %.1f s, mean ratio %.4f" % (len(ev), base["wall_s"], float(np.mean(base["ratio"]))))
    const = constant_predictor(rec) if rec else ConstantPredictor.load()
    res["network"] = evaluate(net, ev, a.iters, log=log, plain=base)
    res["network+default"] = evaluate([net, DefaultInit()], ev, a.iters, log=log, plain=base)
    res["network+default+constant"] = evaluate([net, DefaultInit(), const], ev, a.iters, log=log, plain=base)
    res["constant_per_depth"] = evaluate(const, ev, a.iters, log=log, plain=base)
    res["default_init"] = evaluate(DefaultInit(), ev, a.iters, log=log, plain=base)
    res["default+constant"] = evaluate([DefaultInit(), const], ev, a.iters, log=log, plain=base)
    with open(os.path.join(a.out, "result.json"), "w") as f:
        json.dump(res, f, indent=1)
    log("done in %.0f s" % (time.time() - t0))


def learning_curve(rec, counts, a, log):
    """One network per record prefix (same epochs, learning rate and seed),
    each judged alone and with the per-depth constant of its own prefix on
    the same held-out instances against one plain baseline."""
    import json
    import os
    ev = held_out(a.eval)
    base = plain_baseline(ev, a.iters)
    log("plain baseline: %d instances, %.1f s, mean ratio %.4f" % (len(ev), base["wall_s"], float(np.mean(base["ratio"]))))
    rows = []
    for c in counts:
        sub = subset(rec, c)
        net = Network(seed=a.seed)
        t0 = time.time()
        hist = train(net, sub, a.epochs, lr=a.lr, seed=a.seed, log=None)
        t_train = time.time() - t0
        net.save(os.path.join(a.out, "params_%d.npy" % c))
        table = constant_table(sub)
        with open(os.path.join(a.out, "constant_%d.json" % c), "w") as f:
            json.dump(table, f, indent=1)
        const = ConstantPredictor(table)
        alone = evaluate(net, ev, a.iters, log=None, plain=base)
        pair = evaluate([net, const], ev, a.iters, log=None, plain=base, rescue_iters=6)
        row = {"records": c, "epochs": a.epochs, "train_s": t_train, "final_loss": hist[-1],
               "net_within_1pct": alone["within_1pct_energy"], "net_ratio_raw": alone["mean_ratio_warm_raw"],
               "net_speedup": alone["speedup"], "net_const_within_1pct": pair["within_1pct_energy"],
               "net_const_speedup": pair["speedup"], "net_const_fallbacks": pair["fallbacks"]}
        rows.append(row)
        log("curve %6d records: loss %.5f, train %.0f s | network %.1f%% within 1%%, raw ratio %.4f, %.2fx | "
            "network+constant (rescue 6) %.1f%%, %.2fx, %d fallbacks" % (
                c, row["final_loss"], t_train, 100 * row["net_within_1pct"], row["net_ratio_raw"], row["net_speedup"],
                100 * row["net_const_within_1pct"], row["net_const_speedup"], row["net_const_fallbacks"]))
        with open(os.path.join(a.out, "curve.json"), "w") as f:
            json.dump(rows, f, indent=1)
    return rows


if __name__ == "__main__":
    main()
