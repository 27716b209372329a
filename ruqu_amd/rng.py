"""The reference's seeded RNG over the C ABI: rand 0.8.6 `StdRng`
(`StdRng::seed_from_u64`, state.rs:68; draws `rng.gen::<f64>()`,
state.rs:302 and simulator.rs:187-214), restated in the native library
(ChaCha12 of rand_chacha 0.3.1, rand_core 0.6.4's seed expansion).

A seeded QuantumState / Simulator run takes its measurement and noise draws
from this stream, so seeded outcomes are the reference's own.
"""
import ctypes as C

import numpy as np

from ._native import lib

MASK64 = (1 << 64) - 1


class RqStdRng(C.Structure):
    _fields_ = [("key", C.c_uint32 * 8), ("block", C.c_uint64), ("buf", C.c_uint32 * 64), ("index", C.c_uint32)]


class StdRng:
    """StdRng::seed_from_u64(seed); the seed is taken modulo 2^64 (`as u64`)."""

    def __init__(self, seed):
        self._s = RqStdRng()
        lib().rq_stdrng_seed(C.byref(self._s), int(seed) & MASK64)

    def next_u64(self):
        return int(lib().rq_stdrng_next_u64(C.byref(self._s)))

    def next_u32(self):
        return int(lib().rq_stdrng_next_u32(C.byref(self._s)))

    def random(self, size=None):
        """rng.gen::<f64>(): one float, or an array of `size` of them (in
        stream order), like numpy's Generator.random."""
        n = 1 if size is None else int(np.prod(size))
        out = np.empty(n, dtype=np.float64)
        if n:
            lib().rq_stdrng_next_f64(C.byref(self._s), out.ctypes.data, n)
        if size is None:
            return float(out[0])
        return out.reshape(size)


def shot_draws(base_seed, shots, per_shot):
    """Simulator::run_shots' draws (simulator.rs:122-127): row s is the
    stream of StdRng::seed_from_u64(base_seed + s) (wrapping)."""
    out = np.empty((int(shots), int(per_shot)), dtype=np.float64)
    if out.size:
        lib().rq_stdrng_shot_draws(int(base_seed) & MASK64, int(shots), int(per_shot), out.ctypes.data)
    return out


def state_rng(seed):
    """The draw source of a state: StdRng for a seed, OS entropy otherwise
    (StdRng::from_entropy, state.rs:49)."""
    return np.random.default_rng() if seed is None else StdRng(seed)
