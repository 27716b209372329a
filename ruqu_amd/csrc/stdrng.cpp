// The reference's seeded RNG: rand 0.8.6 `StdRng` (state.rs:68,
// `StdRng::seed_from_u64(seed)`; every measurement and noise draw is
// `rng.gen::<f64>()`, state.rs:302, simulator.rs:187-214).  The crates are
// not vendored in the reference; its Cargo.lock pins rand 0.8.6,
// rand_chacha 0.3.1 and rand_core 0.6.4, whose published algorithms are
// restated here:
//  * StdRng = ChaCha12Rng: the ChaCha block function (Bernstein; RFC 7539
//    section 2.3) with 12 rounds, a 256-bit key, a 64-bit block counter in
//    words 12-13 and a zero 64-bit stream id in words 14-15.  The core fills a
//    buffer of 64 words with 4 consecutive blocks; words are consumed in
//    order.
//  * seed_from_u64 (rand_core 0.6.4): the key is 8 words of a PCG32 sequence
//    (multiplier 6364136223846793005, increment 11634580027462260723; the
//    state advances before each output; output = XSH-RR of the new state).
//  * gen::<f64>() (rand 0.8 Standard): next_u64 >> 11, times 2^-53; next_u64
//    takes the buffer's next two words, low word first (BlockRng).
#include <cstring>

#include "internal.h"

namespace {

inline uint32_t rotl(uint32_t x, int n) { return (x << n) | (x >> (32 - n)); }

inline void quarter(uint32_t* s, int a, int b, int c, int d) {
  s[a] += s[b]; s[d] = rotl(s[d] ^ s[a], 16);
  s[c] += s[d]; s[b] = rotl(s[b] ^ s[c], 12);
  s[a] += s[b]; s[d] = rotl(s[d] ^ s[a], 8);
  s[c] += s[d]; s[b] = rotl(s[b] ^ s[c], 7);
}

// One 64-byte block (16 words) of ChaCha with `rounds` rounds for the input
// state `in` (constants, key, counter / nonce words).
void block(const uint32_t* in, int rounds, uint32_t* out) {
  uint32_t x[16];
  std::memcpy(x, in, sizeof(x));
  for (int r = 0; r < rounds; r += 2) {
    quarter(x, 0, 4, 8, 12); quarter(x, 1, 5, 9, 13); quarter(x, 2, 6, 10, 14); quarter(x, 3, 7, 11, 15);
    quarter(x, 0, 5, 10, 15); quarter(x, 1, 6, 11, 12); quarter(x, 2, 7, 8, 13); quarter(x, 3, 4, 9, 14);
  }
  for (int i = 0; i < 16; ++i) out[i] = x[i] + in[i];
}

void refill(rq_stdrng* r) {
  uint32_t in[16] = {0x61707865u, 0x3320646eu, 0x79622d32u, 0x6b206574u};  // "expand 32-byte k"
  std::memcpy(in + 4, r->key, 32);
  for (int b = 0; b < 4; ++b) {
    const uint64_t ctr = r->block + (uint64_t)b;
    in[12] = (uint32_t)ctr;
    in[13] = (uint32_t)(ctr >> 32);
    in[14] = 0;
    in[15] = 0;
    block(in, 12, r->buf + 16 * b);
  }
  r->block += 4;
}

}  // namespace

extern "C" {

void rq_chacha_block(const uint32_t* in, int rounds, uint32_t* out) { block(in, rounds, out); }

void rq_stdrng_seed(rq_stdrng* r, uint64_t seed) {
  uint64_t st = seed;
  for (int i = 0; i < 8; ++i) {
    st = st * 6364136223846793005ull + 11634580027462260723ull;
    const uint32_t xs = (uint32_t)(((st >> 18) ^ st) >> 27), rot = (uint32_t)(st >> 59);
    r->key[i] = (xs >> rot) | (xs << ((32u - rot) & 31u));
  }
  r->block = 0;
  r->index = 64;  // empty buffer: the first draw generates
  std::memset(r->buf, 0, sizeof(r->buf));
}

uint64_t rq_stdrng_next_u64(rq_stdrng* r) {
  if (r->index < 63) {
    const uint64_t v = ((uint64_t)r->buf[r->index + 1] << 32) | r->buf[r->index];
    r->index += 2;
    return v;
  }
  if (r->index >= 64) {
    refill(r);
    r->index = 2;
    return ((uint64_t)r->buf[1] << 32) | r->buf[0];
  }
  const uint64_t lo = r->buf[63];  // one word left: it is the low half
  refill(r);
  r->index = 1;
  return ((uint64_t)r->buf[0] << 32) | lo;
}

uint32_t rq_stdrng_next_u32(rq_stdrng* r) {
  if (r->index >= 64) {
    refill(r);
    r->index = 0;
  }
  return r->buf[r->index++];
}

void rq_stdrng_next_f64(rq_stdrng* r, double* out, size_t n) {
  for (size_t i = 0; i < n; ++i) out[i] = (double)(rq_stdrng_next_u64(r) >> 11) * (1.0 / 9007199254740992.0);
}

void rq_stdrng_shot_draws(uint64_t base_seed, size_t shots, size_t per_shot, double* out) {
  for (size_t s = 0; s < shots; ++s) {
    rq_stdrng r;
    rq_stdrng_seed(&r, base_seed + (uint64_t)s);
    rq_stdrng_next_f64(&r, out + s * per_shot, per_shot);
  }
}

}  // extern "C"
