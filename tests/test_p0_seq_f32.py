"""The complex-f32 measurement sums P(q=0) the way QuantumStateF32::measure
does (mixed_precision.rs:480-486): a sequential chain of f32 additions in
index order.  The device kernel (k_p0_seq_f32) evaluates that chain in
chunks: while the running sum s sits in one binade [2^E, 2^(E+1)), each step
rounds s + x to s + round(x / u) u with u = 2^(E-23), so a chunk without
halfway cases whose total stays in the binade is one integer sum; every other
chunk is added term by term.

These CPU tests check that claim on a numpy model of the chunk rule against
numpy's sequential float32 accumulate (np.add.accumulate rounds every partial
sum), on inputs built to hit each branch: ties at u/2, binade crossings
inside a chunk, zero and subnormal sums, and the stall where terms fall below
half an ulp of the sum.  The device kernel itself is compared with the
oracle's sequential sum in tests/test_gpu_parity.py."""
import numpy as np
import pytest

FLT_MIN = np.float32(1.17549435e-38)


def sequential_f32(x):
    return np.add.accumulate(np.asarray(x, dtype=np.float32), dtype=np.float32)[-1] if len(x) else np.float32(0)


def chunked_f32(x, chunk):
    """The kernel's rule (chunk = threads x terms per thread)."""
    x = np.asarray(x, dtype=np.float32)
    s = np.float32(0.0)
    stats = {"parallel": 0, "serial": 0}
    for base in range(0, len(x), chunk):
        c = x[base:base + chunk]
        if s >= FLT_MIN:
            _, e2 = np.frexp(np.float64(s))
            E = int(e2) - 1
            y = np.ldexp(c.astype(np.float64), 23 - E)
            bad = bool(np.any(~(y < 2.0 ** 25)))
            f = np.floor(np.where(y < 2.0 ** 25, y, 0.0))
            d = np.where(y < 2.0 ** 25, y, 0.0) - f
            bad |= bool(np.any(d == 0.5))
            tot = int(np.sum(f.astype(np.int64))) + int(np.count_nonzero(d > 0.5))
            su = int(np.ldexp(np.float64(s), 23 - E))
            assert 2 ** 23 <= su < 2 ** 24
            if not bad and su + tot < 2 ** 24:
                s = np.float32(np.ldexp(np.float64(su + tot), E - 23))
                stats["parallel"] += 1
                continue
        stats["serial"] += 1
        for v in c:
            s = np.float32(s + v)
    return s, stats


def _norm_sq_f32(a):
    re, im = a.real.astype(np.float32), a.imag.astype(np.float32)
    return re * re + im * im


@pytest.mark.parametrize("n,seed", [(12, 1), (16, 2), (18, 3)])
def test_random_states(n, seed):
    rng = np.random.default_rng(seed)
    a = (rng.normal(size=1 << n) + 1j * rng.normal(size=1 << n))
    a /= np.linalg.norm(a)
    x = _norm_sq_f32(a.astype(np.complex64))
    for chunk in (64, 1024, 16384):
        got, st = chunked_f32(x, chunk)
        assert got == sequential_f32(x)
        if len(x) >= 8 * chunk:
            assert st["parallel"] > st["serial"]


def test_stall_regime():
    """2^25 equal terms of 0.9 * 2^-26 (what a 26-qubit state of equal
    magnitudes would give): the true sum is 0.45, but once the sum reaches
    0.25 every term is below half an ulp and the f32 chain stops there.  The
    chunk rule must stall at exactly the same value."""
    x = np.full(1 << 25, np.float32(0.9 * 2.0 ** -26))
    want = sequential_f32(x)
    assert want == np.float32(0.25)  # the reference's f32 sum, far from 0.45
    got, st = chunked_f32(x, 16384)
    assert got == want
    assert st["parallel"] > 100 * st["serial"]


def test_ties_fall_back_to_serial():
    """Terms at exactly half an ulp of the running sum: round-to-even depends
    on s, so the chunk rule must not take the integer sum."""
    s0 = np.float32(1.0)
    u = np.float32(2.0 ** -23)
    x = np.concatenate([[s0], np.full(100, u / 2), np.full(37, np.float32(3) * u / 2), [u * np.float32(0.75)]])
    x = x.astype(np.float32)
    got, st = chunked_f32(x, 32)
    assert got == sequential_f32(x)
    assert st["serial"] >= 2


def test_binade_crossings_and_subnormals():
    rng = np.random.default_rng(9)
    parts = [np.zeros(50, np.float32),
             np.full(40, np.float32(1e-45)),            # subnormal terms from a zero sum
             rng.uniform(0, 1e-39, 300).astype(np.float32),
             np.geomspace(1e-30, 1e-2, 500).astype(np.float32),  # sum crosses many binades
             rng.uniform(0, 1e-3, 4000).astype(np.float32),
             np.float32([0.3, 0.6, 0.2]),
             rng.uniform(0, 1e-9, 5000).astype(np.float32)]
    x = np.concatenate(parts)
    for chunk in (16, 128, 1024):
        got, st = chunked_f32(x, chunk)
        assert got == sequential_f32(x), chunk
        assert st["parallel"] > 0 and st["serial"] > 0
