"""Host side of the warm-start records: compact record files (no graphs)
regenerate the same graphs as the full format, and prefixes/constants are
taken over the right records.  No device needed."""
import numpy as np

from ruqu_amd import warmstart as W


def fake_records(count, start):
    rng = np.random.default_rng(start + 1)
    insts = W.instances(count, start=start)
    return {"graphs": [(n, e, w) for n, e, w, _ in insts], "p": np.array([p for *_, p in insts]),
            "gammas": [rng.uniform(0, 1, p) for *_, p in insts], "betas": [rng.uniform(0, 1, p) for *_, p in insts],
            "best": rng.uniform(1, 10, count)}


def test_compact_records_round_trip(tmp_path):
    a, b = fake_records(40, 0), fake_records(25, 40)
    W.save_records_compact(tmp_path / "a.npz", a, 0)
    W.save_records_compact(tmp_path / "b.npz", b, 40)
    W.save_records(tmp_path / "full.npz", a)
    got = W.load_records_compact([tmp_path / "a.npz", tmp_path / "b.npz"])
    full = W.load_records(tmp_path / "full.npz")
    assert len(got["p"]) == 65
    norm = [(n, [tuple(x) for x in e], w) for n, e, w in full["graphs"]]
    assert got["graphs"][:40] == norm
    assert got["graphs"] == [(n, e, w) for n, e, w, _ in W.instances(65)]
    for i in range(65):
        src = a if i < 40 else b
        j = i if i < 40 else i - 40
        assert np.array_equal(got["gammas"][i], src["gammas"][j])
        assert np.array_equal(got["betas"][i], src["betas"][j])
    assert np.array_equal(got["best"], np.concatenate([a["best"], b["best"]]))


def test_compact_records_reject_a_wrong_start(tmp_path):
    a = fake_records(30, 0)
    W.save_records_compact(tmp_path / "a.npz", a, 0)
    with np.load(tmp_path / "a.npz") as z:
        arrs = {k: z[k] for k in z.files}
    arrs["start"] = np.int64(1)  # the depths no longer match the regenerated instances
    np.savez_compressed(tmp_path / "shifted.npz", **arrs)
    try:
        W.load_records_compact([tmp_path / "shifted.npz"])
    except ValueError:
        return
    assert False, "a shifted record file was accepted"


def test_subset_and_constant_predictor():
    rec = fake_records(50, 0)
    sub = W.subset(rec, 20)
    assert len(sub["graphs"]) == 20 and np.array_equal(sub["p"], rec["p"][:20])
    const = W.constant_predictor(sub)
    for p in set(int(x) for x in sub["p"]):
        sel = [i for i in range(20) if sub["p"][i] == p]
        g, bt = const.predict(None, [p])[0]
        assert np.allclose(g, np.mean([sub["gammas"][i] for i in sel], axis=0))
        assert np.allclose(bt, np.mean([sub["betas"][i] for i in sel], axis=0))
