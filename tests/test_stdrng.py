"""The seeded RNG (rand 0.8.6 StdRng, state.rs:68) — the oracle's
restatement pinned by published vectors, and the native library's version
checked against it.  CPU only (the RNG is host code)."""
import ctypes as C

import numpy as np
import pytest

from oracle.stdrng_oracle import SIGMA, StdRngOracle, chacha_block, pcg32_reference
from ruqu_amd._native import lib
from ruqu_amd.rng import StdRng, shot_draws


def words(hexbytes):
    b = bytes.fromhex(hexbytes.replace(" ", ""))
    return [int.from_bytes(b[i:i + 4], "little") for i in range(0, len(b), 4)]


KEY0 = [0] * 8
# RFC 7539 appendix A.1, test vectors 1 and 2 (ChaCha20, zero key and nonce).
A1_1 = words("76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
             "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
A1_2 = words("9f07e7be5551387a98ba977c732d080dcb0f29a048e3656912c6533e32ee7aed"
             "29b721769ce64e43d57133b074d839d531ed1f28510afb45ace10a1f4b794d6f")
# RFC 7539 section 2.3.2: key 00..1f, counter 1, nonce 00000009 0000004a 00000000.
KEY_SEQ = words(bytes(range(32)).hex())
S232_IN = list(SIGMA) + KEY_SEQ + [1, 0x09000000, 0x4A000000, 0]
S232 = [0xE4E7F110, 0x15593BD1, 0x1FDD0F50, 0xC47120A3, 0xC7F4D1C7, 0x0368C033, 0x9AAA2204, 0x4E6CD4C3,
        0x466482D2, 0x09AA9F07, 0x05D7C214, 0xA2028BD9, 0xD19C12B5, 0xB94E16DE, 0xE883D0CB, 0x4E3C50A2]
VECTORS = [(list(SIGMA) + KEY0 + [0, 0, 0, 0], A1_1), (list(SIGMA) + KEY0 + [1, 0, 0, 0], A1_2), (S232_IN, S232)]


def native_block(state, rounds):
    inp = (C.c_uint32 * 16)(*state)
    out = (C.c_uint32 * 16)()
    lib().rq_chacha_block(inp, rounds, out)
    return list(out)


@pytest.mark.parametrize("k", range(3))
def test_chacha20_rfc7539_vectors(k):
    state, want = VECTORS[k]
    assert chacha_block(state, 20) == want
    assert native_block(state, 20) == want


def test_pcg32_reference_demo_output():
    # pcg32-demo, pcg32_srandom(42, 54): fixes the multiplier and XSH-RR.
    assert pcg32_reference(42, 54, 6) == [0xA15C02B7, 0x7B47F409, 0xBA1D3330, 0x83D2F293, 0xBFA4784B, 0xCBED606E]


def test_chacha12_differs_from_chacha20_only_in_rounds():
    state = VECTORS[2][0]
    assert chacha_block(state, 12) == native_block(state, 12) != chacha_block(state, 20)


@pytest.mark.parametrize("seed", [0, 1, 42, 12345, (1 << 64) - 1])
def test_native_stream_equals_oracle(seed):
    o, r = StdRngOracle(seed), StdRng(seed)
    # 150 u64 = 300 words: several refills of the 64-word buffer
    assert [r.next_u64() for _ in range(150)] == [o.next_u64() for _ in range(150)]
    # an odd number of u32 words, then u64s across the buffer's last word
    assert [r.next_u32() for _ in range(33)] == [o.next_u32() for _ in range(33)]
    assert [r.next_u64() for _ in range(40)] == [o.next_u64() for _ in range(40)]
    assert list(r.random(100)) == o.random(100)


def test_f64_draws_are_53_bit_fractions():
    d = StdRng(7).random(20000)
    assert d.min() >= 0.0 and d.max() < 1.0
    assert np.all(np.floor(d * 2.0 ** 53) == d * 2.0 ** 53)
    assert abs(d.mean() - 0.5) < 0.01


def test_seed_is_taken_modulo_2_64():
    assert list(StdRng(-1).random(5)) == list(StdRng((1 << 64) - 1).random(5))


def test_shot_draws_are_per_shot_streams():  # simulator.rs:122-127
    d = shot_draws(42, 5, 7)
    for s in range(5):
        assert list(d[s]) == StdRngOracle(42 + s).random(7)
    w = shot_draws((1 << 64) - 2, 3, 2)  # wrapping_add
    assert list(w[2]) == StdRngOracle(0).random(2)
