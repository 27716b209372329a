// ORACLE — test infrastructure only (CPU baseline leg).
//
// Restatement of ruqu-core's `simd.rs` kernels, the "SIMD multi-threaded" CPU
// path that the reference ships as free functions on &mut [Complex]:
//   * apply_single_qubit_gate_simd (AVX2, simd.rs:109-181, helper :200-233)
//   * apply_single_qubit_gate_parallel (rayon chunks, simd.rs:270-309)
//   * apply_two_qubit_gate_parallel (materialised base list, simd.rs:319-377)
//   * the *_best dispatch (simd.rs:393-445): parallel when len >= 65 536,
//     else AVX2 for 1q, scalar for 2q.
// rayon's pool becomes a fixed set of std::threads (one per requested core)
// that split the chunk list statically.  Built with -O3 -mavx2
// -ffp-contract=off; the AVX2 path performs the same IEEE operations in the
// same order as the scalar butterfly (mul, mul by +-1, add), so results are
// bit-identical to the scalar restatement, which the tests check.

#include <immintrin.h>

#include <cstdint>
#include <cstring>
#include <thread>
#include <vector>

namespace {

struct C64 {
  double re, im;
};
inline C64 operator+(C64 a, C64 b) { return C64{a.re + b.re, a.im + b.im}; }
inline C64 operator*(C64 a, C64 b) {
  return C64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}

const size_t PARALLEL_THRESHOLD = 65536;  // simd.rs:25

// simd.rs:37-56
void scalar_1q(C64* a, size_t n, uint32_t q, const C64 m[2][2]) {
  size_t step = (size_t)1 << q;
  for (size_t block = 0; block < n; block += step << 1)
    for (size_t i = block; i < block + step; ++i) {
      size_t j = i + step;
      C64 x = a[i], y = a[j];
      a[i] = m[0][0] * x + m[0][1] * y;
      a[j] = m[1][0] * x + m[1][1] * y;
    }
}

// simd.rs:60-91
void scalar_2q(C64* a, size_t n, uint32_t q1, uint32_t q2, const C64 m[4][4]) {
  size_t b1 = (size_t)1 << q1, b2 = (size_t)1 << q2;
  for (size_t base = 0; base < n; ++base) {
    if ((base & b1) || (base & b2)) continue;
    size_t idx[4] = {base, base | b2, base | b1, base | b1 | b2};
    C64 v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
    for (int r = 0; r < 4; ++r)
      a[idx[r]] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2] + m[r][3] * v[3];
  }
}

// simd.rs:200-233
inline __m256d cmul_add(__m256d a, __m256d ar, __m256d ai, __m256d b, __m256d br, __m256d bi, __m256d neg) {
  __m256d as = _mm256_permute_pd(a, 0b0101);
  __m256d ra = _mm256_add_pd(_mm256_mul_pd(ar, a), _mm256_mul_pd(_mm256_mul_pd(ai, as), neg));
  __m256d bs = _mm256_permute_pd(b, 0b0101);
  __m256d rb = _mm256_add_pd(_mm256_mul_pd(br, b), _mm256_mul_pd(_mm256_mul_pd(bi, bs), neg));
  return _mm256_add_pd(ra, rb);
}

// simd.rs:109-181
void avx2_1q(C64* a, size_t n, uint32_t q, const C64 m[2][2]) {
  size_t step = (size_t)1 << q;
  if (step < 2) { scalar_1q(a, n, q, m); return; }
  __m256d m00r = _mm256_set1_pd(m[0][0].re), m00i = _mm256_set1_pd(m[0][0].im);
  __m256d m01r = _mm256_set1_pd(m[0][1].re), m01i = _mm256_set1_pd(m[0][1].im);
  __m256d m10r = _mm256_set1_pd(m[1][0].re), m10i = _mm256_set1_pd(m[1][0].im);
  __m256d m11r = _mm256_set1_pd(m[1][1].re), m11i = _mm256_set1_pd(m[1][1].im);
  __m256d neg = _mm256_set_pd(-1.0, 1.0, -1.0, 1.0);
  for (size_t block = 0; block < n; block += step << 1) {
    size_t i = block;
    for (; i + 1 < block + step; i += 2) {
      size_t j = i + step;
      __m256d av = _mm256_loadu_pd((const double*)&a[i]);
      __m256d bv = _mm256_loadu_pd((const double*)&a[j]);
      __m256d oi = cmul_add(av, m00r, m00i, bv, m01r, m01i, neg);
      __m256d oj = cmul_add(av, m10r, m10i, bv, m11r, m11i, neg);
      _mm256_storeu_pd((double*)&a[i], oi);
      _mm256_storeu_pd((double*)&a[j], oj);
    }
    if (step & 1) {
      size_t ii = block + step - 1, j = ii + step;
      C64 x = a[ii], y = a[j];
      a[ii] = m[0][0] * x + m[0][1] * y;
      a[j] = m[1][0] * x + m[1][1] * y;
    }
  }
}

template <class F>
void run_threads(int nthreads, size_t items, F f) {
  if (nthreads <= 1 || items <= 1) { f(0, items); return; }
  std::vector<std::thread> th;
  size_t per = (items + nthreads - 1) / nthreads;
  for (int t = 0; t < nthreads; ++t) {
    size_t lo = t * per, hi = lo + per < items ? lo + per : items;
    if (lo >= hi) break;
    th.emplace_back([=] { f(lo, hi); });
  }
  for (auto& x : th) x.join();
}

// simd.rs:270-309 — chunks of whole butterfly blocks, >= 4096 amplitudes,
// scalar butterflies inside each chunk.
void parallel_1q(C64* a, size_t n, uint32_t q, const C64 m[2][2], int nthreads) {
  if (n < PARALLEL_THRESHOLD) { scalar_1q(a, n, q, m); return; }
  size_t step = (size_t)1 << q, block = step << 1;
  size_t min_chunk = block > 4096 ? block : 4096;
  size_t chunk = ((min_chunk + block - 1) / block) * block;
  size_t nchunks = (n + chunk - 1) / chunk;
  run_threads(nthreads, nchunks, [&](size_t lo, size_t hi) {
    for (size_t c = lo; c < hi; ++c) {
      C64* ch = a + c * chunk;
      size_t len = (c + 1) * chunk <= n ? chunk : n - c * chunk;
      for (size_t bs = 0; bs + block <= len; bs += block)
        for (size_t i = bs; i < bs + step; ++i) {
          size_t j = i + step;
          C64 x = ch[i], y = ch[j];
          ch[i] = m[0][0] * x + m[0][1] * y;
          ch[j] = m[1][0] * x + m[1][1] * y;
        }
    }
  });
}

// simd.rs:319-377 — materialise every base index, then split the list.
void parallel_2q(C64* a, size_t n, uint32_t q1, uint32_t q2, const C64 m[4][4], int nthreads) {
  if (n < PARALLEL_THRESHOLD) { scalar_2q(a, n, q1, q2, m); return; }
  size_t b1 = (size_t)1 << q1, b2 = (size_t)1 << q2;
  std::vector<size_t> bases;
  bases.reserve(n / 4);
  for (size_t base = 0; base < n; ++base)
    if ((base & b1) == 0 && (base & b2) == 0) bases.push_back(base);
  run_threads(nthreads, bases.size(), [&](size_t lo, size_t hi) {
    for (size_t k = lo; k < hi; ++k) {
      size_t base = bases[k];
      size_t idx[4] = {base, base | b2, base | b1, base | b1 | b2};
      C64 v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
      for (int r = 0; r < 4; ++r)
        a[idx[r]] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2] + m[r][3] * v[3];
    }
  });
}

}  // namespace

extern "C" {

// simd.rs:393-422 — apply_single_qubit_gate_best with features simd+parallel.
void orc_simd_apply_1q_best(double* amps, uint32_t nq, uint32_t q, const double m[8], int nthreads) {
  C64 mm[2][2]; memcpy(mm, m, sizeof(mm));
  size_t n = (size_t)1 << nq;
  if (n >= PARALLEL_THRESHOLD) parallel_1q((C64*)amps, n, q, mm, nthreads);
  else avx2_1q((C64*)amps, n, q, mm);
}

// simd.rs:428-445
void orc_simd_apply_2q_best(double* amps, uint32_t nq, uint32_t q1, uint32_t q2, const double m[32], int nthreads) {
  C64 mm[4][4]; memcpy(mm, m, sizeof(mm));
  size_t n = (size_t)1 << nq;
  if (n >= PARALLEL_THRESHOLD) parallel_2q((C64*)amps, n, q1, q2, mm, nthreads);
  else scalar_2q((C64*)amps, n, q1, q2, mm);
}

// The AVX2 kernel on its own (simd.rs:109-181), for the bit-identity test.
void orc_simd_apply_1q_avx2(double* amps, uint32_t nq, uint32_t q, const double m[8]) {
  C64 mm[2][2]; memcpy(mm, m, sizeof(mm));
  avx2_1q((C64*)amps, (size_t)1 << nq, q, mm);
}

}  // extern "C"
