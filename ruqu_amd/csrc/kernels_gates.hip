// Exact gate kernels: one HBM pass per gate, the reference's arithmetic.
//
// Every kernel here evaluates the same IEEE expression as ruqu-core's scalar
// loops (state.rs:214-276, mixed_precision.rs:396-458): complex products
// m.re*a.re - m.im*a.im / m.re*a.im + m.im*a.re and left-to-right row sums,
// with FP contraction disabled, so f64 and f32 results equal the reference's
// bit for bit (up to the sign of exact zeros, which the reference's dense
// 4x4 sums can flip).  Specialised 2q kernels (CNOT, CZ, SWAP, diagonal) skip
// only the products by exact zeros and ones.
//
// Layout: interleaved (re, im); f32 kernels move 16 B per load (two
// amplitudes, a float4) wherever both amplitudes share the pair structure.
#include "internal.h"

#pragma clang fp contract(off)

namespace rq {
namespace {

constexpr int BLOCK = 256;

template <class T>
struct Cx {
  T re, im;
};

template <class T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> m, Cx<T> a) {
  return Cx<T>{m.re * a.re - m.im * a.im, m.re * a.im + m.im * a.re};
}
template <class T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) {
  return Cx<T>{a.re + b.re, a.im + b.im};
}

template <class T>
struct M2 {
  Cx<T> m00, m01, m10, m11;
};
template <class T>
struct M4 {
  Cx<T> m[16];
};

// HBM access with an optional non-temporal hint, through the 16-byte native
// vector type (float4 and double2 are both 16 bytes).
typedef float nvec4 __attribute__((ext_vector_type(4)));
template <bool NT, class V>
__device__ __forceinline__ V ldg(const V* p) {
  static_assert(sizeof(V) == 16, "16-byte vectors only");
  if constexpr (NT) {
    nvec4 r = __builtin_nontemporal_load(reinterpret_cast<const nvec4*>(p));
    return __builtin_bit_cast(V, r);
  } else {
    return *p;
  }
}
template <bool NT, class V>
__device__ __forceinline__ void stg(V* p, V v) {
  static_assert(sizeof(V) == 16, "16-byte vectors only");
  if constexpr (NT) __builtin_nontemporal_store(__builtin_bit_cast(nvec4, v), reinterpret_cast<nvec4*>(p));
  else *p = v;
}

// Workgroup order of the two-stream kernels.  Workgroups are dispatched to the
// 8 XCDs in turn; with chunk C > 0 XCD x takes C consecutive pair blocks at a
// time (chunk k of XCD x is blocks (8k + x) C .. (8k + x) C + C - 1), so the
// blocks in flight on one XCD are contiguous in memory.  A bijection on
// [0, nb) when nb is a multiple of 8 C; the host passes C = 0 otherwise.
__device__ __forceinline__ uint64_t xcd_block(uint32_t C) {
  const uint64_t b = flat_block();
  if (C == 0) return b;
  const uint64_t x = b & 7u, r = b >> 3;
  return ((r / C) * 8u + x) * C + (r % C);
}

// A zero computed from a loaded value, opaque to the compiler.  The second
// stream's addresses of the two-stream kernels are offset by it, so each
// thread issues its partner loads only when its first loads have returned:
// the two loads of a pair (2^q amplitudes apart) are then not in flight
// together.  Issued together, pairs 16 MB, 256 MB and 4 GB apart ran at
// 5.7-5.8 TB/s on MI355X; issued in turn, with the chunks below, every
// distance from 1 KB to 4 GB runs at 6.46-6.84 TB/s.
__device__ __forceinline__ uint32_t zero_after(uint32_t v) {
  asm volatile("v_and_b32 %0, 0, %0" : "+v"(v));
  return v;
}

// Insert a zero bit at position q of p.
__device__ __forceinline__ uint64_t ins0(uint64_t p, uint32_t q) {
  uint64_t lo = p & ((1ull << q) - 1);
  return ((p >> q) << (q + 1)) | lo;
}

// ---- 1q, f32, q >= 1: each group = two adjacent pairs (float4 each side) --
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_1q_f32_v2(float4* __restrict__ a, uint32_t q, uint64_t ngroups,
                                                     M2<float> m, uint32_t chunk) {
  const uint64_t g0 = xcd_block(chunk) * U * BLOCK + threadIdx.x;
  float4 x[U], y[U];
  uint64_t ii[U], jj[U];
  uint32_t z = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < ngroups) {
      uint64_t i = ins0(2 * g, q);
      ii[k] = i >> 1;
      jj[k] = (i + (1ull << q)) >> 1;
      x[k] = ldg<NT>(a + ii[k]);
      z = __float_as_uint(x[k].x);
    }
  }
  z = zero_after(z);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < ngroups) y[k] = ldg<NT>(a + jj[k] + z);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < ngroups) {
      Cx<float> a0{x[k].x, x[k].y}, a1{x[k].z, x[k].w};
      Cx<float> b0{y[k].x, y[k].y}, b1{y[k].z, y[k].w};
      Cx<float> o0 = cadd(cmul(m.m00, a0), cmul(m.m01, b0));
      Cx<float> o1 = cadd(cmul(m.m00, a1), cmul(m.m01, b1));
      Cx<float> p0 = cadd(cmul(m.m10, a0), cmul(m.m11, b0));
      Cx<float> p1 = cadd(cmul(m.m10, a1), cmul(m.m11, b1));
      stg<NT>(a + ii[k], make_float4(o0.re, o0.im, o1.re, o1.im));
      stg<NT>(a + jj[k], make_float4(p0.re, p0.im, p1.re, p1.im));
    }
  }
}

// ---- 1q, f32, 1 <= q <= 6: partner vectors in the same wave ---------------
// Each lane reads one float4 (amplitudes 2v, 2v + 1) of a contiguous run, so
// every load and store of a wave covers 1 KB; the pair partner of vector v is
// v ^ 2^(q-1), another lane of the same wave, whose value comes through a
// lane shuffle.  A lane on the |0> side writes m00 a + m01 b, one on the |1>
// side m10 a + m11 b: the same operations in the same order as k_1q_f32_v2.
// Needs nvec (= 2^(n-1)) to be a multiple of 64: whole waves in or out.
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_1q_f32_lx(float4* __restrict__ a, uint32_t q, uint64_t nvec,
                                                     M2<float> m) {
  const uint64_t g0 = flat_block() * U * BLOCK + threadIdx.x;
  const int d = 1 << (q - 1);
  const bool one = (threadIdx.x >> (q - 1)) & 1u;  // this vector holds the amplitudes with bit q = 1
  const Cx<float> m0 = one ? m.m10 : m.m00, m1 = one ? m.m11 : m.m01;
  float4 x[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t v = g0 + (uint64_t)k * BLOCK;
    if (v < nvec) x[k] = ldg<NT>(a + v);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t v = g0 + (uint64_t)k * BLOCK;
    if (v < nvec) {
      const float4 y = make_float4(__shfl_xor(x[k].x, d), __shfl_xor(x[k].y, d), __shfl_xor(x[k].z, d),
                                   __shfl_xor(x[k].w, d));
      const float4 lo = one ? y : x[k], hi = one ? x[k] : y;  // |0> side, |1> side
      const Cx<float> a0{lo.x, lo.y}, a1{lo.z, lo.w}, b0{hi.x, hi.y}, b1{hi.z, hi.w};
      const Cx<float> o0 = cadd(cmul(m0, a0), cmul(m1, b0));
      const Cx<float> o1 = cadd(cmul(m0, a1), cmul(m1, b1));
      stg<NT>(a + v, make_float4(o0.re, o0.im, o1.re, o1.im));
    }
  }
}

// ---- 1q, f32, q == 0: one float4 holds the pair (a, b) --------------------
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_1q_f32_q0(float4* __restrict__ a, uint64_t npairs, M2<float> m) {
  const uint64_t g0 = flat_block() * U * BLOCK + threadIdx.x;
  float4 x[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < npairs) x[k] = ldg<NT>(a + g);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < npairs) {
      Cx<float> av{x[k].x, x[k].y}, bv{x[k].z, x[k].w};
      Cx<float> o = cadd(cmul(m.m00, av), cmul(m.m01, bv));
      Cx<float> p = cadd(cmul(m.m10, av), cmul(m.m11, bv));
      stg<NT>(a + g, make_float4(o.re, o.im, p.re, p.im));
    }
  }
}

// ---- 1q, f64: one double2 per amplitude -----------------------------------
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_1q_f64(double2* __restrict__ a, uint32_t q, uint64_t npairs,
                                                  M2<double> m, uint32_t chunk) {
  const uint64_t g0 = xcd_block(chunk) * U * BLOCK + threadIdx.x;
  double2 x[U], y[U];
  uint64_t ii[U];
  uint32_t z = 0;
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < npairs) {
      ii[k] = ins0(g, q);
      x[k] = ldg<NT>(a + ii[k]);
      z = (uint32_t)__double_as_longlong(x[k].x);
    }
  }
  z = zero_after(z);
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < npairs) y[k] = ldg<NT>(a + ii[k] + (1ull << q) + z);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < npairs) {
      Cx<double> av{x[k].x, x[k].y}, bv{y[k].x, y[k].y};
      Cx<double> o = cadd(cmul(m.m00, av), cmul(m.m01, bv));
      Cx<double> p = cadd(cmul(m.m10, av), cmul(m.m11, bv));
      stg<NT>(a + ii[k], make_double2(o.re, o.im));
      stg<NT>(a + ii[k] + (1ull << q), make_double2(p.re, p.im));
    }
  }
}

// ---- 1q, f64, q <= 5: partner amplitudes in the same wave -----------------
// Lane l reads amplitude v = (wave base) + l; its pair partner v ^ 2^q is
// lane l ^ 2^q, read through lane shuffles, so every load and store of a wave
// covers 1 KB.  Same operations in the same order as k_1q_f64.  Needs n to
// be a multiple of 64 (whole waves in or out).
__device__ __forceinline__ double shfl_xor_f64(double x, int d) {
  const unsigned long long u = __double_as_longlong(x);
  const int lo = __shfl_xor((int)(unsigned)u, d), hi = __shfl_xor((int)(unsigned)(u >> 32), d);
  return __longlong_as_double(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_1q_f64_lx(double2* __restrict__ a, uint32_t q, uint64_t n,
                                                     M2<double> m) {
  const uint64_t g0 = flat_block() * U * BLOCK + threadIdx.x;
  const int d = 1 << q;
  const bool one = (threadIdx.x >> q) & 1u;  // this amplitude has bit q = 1
  const Cx<double> m0 = one ? m.m10 : m.m00, m1 = one ? m.m11 : m.m01;
  double2 x[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t v = g0 + (uint64_t)k * BLOCK;
    if (v < n) x[k] = ldg<NT>(a + v);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    const uint64_t v = g0 + (uint64_t)k * BLOCK;
    if (v < n) {
      const Cx<double> own{x[k].x, x[k].y}, oth{shfl_xor_f64(x[k].x, d), shfl_xor_f64(x[k].y, d)};
      const Cx<double> av = one ? oth : own, bv = one ? own : oth;
      const Cx<double> o = cadd(cmul(m0, av), cmul(m1, bv));
      stg<NT>(a + v, make_double2(o.re, o.im));
    }
  }
}

// ---- dense 2q (state.rs:238-276): one quad per thread ---------------------
template <class T, class V>
__global__ __launch_bounds__(BLOCK) void k_2q_dense(V* __restrict__ a, uint32_t q1, uint32_t q2, uint64_t nquads,
                                                    M4<T> m) {
  uint64_t g = flat_block() * BLOCK + threadIdx.x;
  if (g >= nquads) return;
  uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  uint64_t base = ins0(ins0(g, lo), hi);
  uint64_t b1 = 1ull << q1, b2 = 1ull << q2;
  uint64_t idx[4] = {base, base | b2, base | b1, base | b1 | b2};
  Cx<T> v[4];
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    V t = a[idx[r]];
    v[r] = Cx<T>{t.x, t.y};
  }
#pragma unroll
  for (int r = 0; r < 4; ++r) {
    Cx<T> s = cadd(cadd(cadd(cmul(m.m[r * 4 + 0], v[0]), cmul(m.m[r * 4 + 1], v[1])), cmul(m.m[r * 4 + 2], v[2])),
                   cmul(m.m[r * 4 + 3], v[3]));
    V o;
    o.x = s.re;
    o.y = s.im;
    a[idx[r]] = o;
  }
}

// ---- CNOT(c, t): swap the t=0 / t=1 amplitudes where c=1 ------------------
template <class V>
__global__ __launch_bounds__(BLOCK) void k_cnot(V* __restrict__ a, uint32_t c, uint32_t t, uint64_t nquads) {
  uint64_t g = flat_block() * BLOCK + threadIdx.x;
  if (g >= nquads) return;
  uint32_t lo = c < t ? c : t, hi = c < t ? t : c;
  uint64_t base = ins0(ins0(g, lo), hi) | (1ull << c);
  V x = a[base], y = a[base | (1ull << t)];
  a[base] = y;
  a[base | (1ull << t)] = x;
}

// ---- SWAP(q1, q2): exchange |01> and |10> ---------------------------------
template <class V>
__global__ __launch_bounds__(BLOCK) void k_swap(V* __restrict__ a, uint32_t q1, uint32_t q2, uint64_t nquads) {
  uint64_t g = flat_block() * BLOCK + threadIdx.x;
  if (g >= nquads) return;
  uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  uint64_t base = ins0(ins0(g, lo), hi);
  uint64_t i = base | (1ull << q2), j = base | (1ull << q1);
  V x = a[i], y = a[j];
  a[i] = y;
  a[j] = x;
}

// ---- CZ: (-1 + 0i) * a on |11> --------------------------------------------
template <class T, class V>
__global__ __launch_bounds__(BLOCK) void k_cz(V* __restrict__ a, uint32_t q1, uint32_t q2, uint64_t nquads) {
  uint64_t g = flat_block() * BLOCK + threadIdx.x;
  if (g >= nquads) return;
  uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  uint64_t i = ins0(ins0(g, lo), hi) | (1ull << q1) | (1ull << q2);
  V x = a[i];
  Cx<T> r = cmul(Cx<T>{T(-1), T(0)}, Cx<T>{x.x, x.y});
  x.x = r.re;
  x.y = r.im;
  a[i] = x;
}

// ---- CZ / CNOT as one streaming pass -------------------------------------
// When both CZ qubits (a CNOT control) lie within the 128-byte line, every
// line holds amplitudes the gate changes, so the gate costs a full pass
// whichever way it is done; these kernels make it one: every lane reads and writes one contiguous 16-byte
// vector (complex f32: amplitudes 2v, 2v + 1; complex f64: amplitude v).  A
// CNOT partner vector is lane l ^ 2^(t - log2(amplitudes per vector)) of the
// same wave (lane shuffle).  Same operations as k_cz / k_cnot (a product by
// -1 + 0i; exact moves).  With `skip`, CZ lanes whose vector does not change
// neither load nor store: that pays when the changed runs are >= 32 bytes
// (a quarter of each line is written); shorter partial writes cost more than
// rewriting the line.  Need the vector count to be a multiple of 64.
template <class T>
__device__ __forceinline__ Cx<T> cz_neg(Cx<T> x) {
  return cmul(Cx<T>{T(-1), T(0)}, x);
}
__global__ __launch_bounds__(BLOCK) void k_cz_full_f32(float4* __restrict__ a, uint32_t q1, uint32_t q2,
                                                       uint64_t nvec, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  const uint64_t m = (1ull << q1) | (1ull << q2), i0 = 2 * v;
  if (v >= nvec || (skip && (i0 & m) != m && ((i0 + 1) & m) != m)) return;
  float4 x = a[v];
  if ((i0 & m) == m) {
    const Cx<float> r = cz_neg(Cx<float>{x.x, x.y});
    x.x = r.re;
    x.y = r.im;
  }
  if (((i0 + 1) & m) == m) {
    const Cx<float> r = cz_neg(Cx<float>{x.z, x.w});
    x.z = r.re;
    x.w = r.im;
  }
  a[v] = x;
}
__global__ __launch_bounds__(BLOCK) void k_cz_full_f64(double2* __restrict__ a, uint32_t q1, uint32_t q2,
                                                       uint64_t n, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  const uint64_t m = (1ull << q1) | (1ull << q2);
  const bool chg = (v & m) == m;
  if (v >= n || (skip && !chg)) return;
  const double2 x = a[v];
  const Cx<double> r = cz_neg(Cx<double>{x.x, x.y});
  a[v] = chg ? make_double2(r.re, r.im) : x;
}
// complex f32, t <= 6 (t = 0: the partner is the other half of the vector).
// With `skip` (c >= 1), lanes whose vector has the control bit clear neither
// load nor store: their partners have it clear too, so nothing reads the
// value they pass through the shuffle.
__global__ __launch_bounds__(BLOCK) void k_cnot_full_f32(float4* __restrict__ a, uint32_t c, uint32_t t,
                                                         uint64_t nvec, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  if (v >= nvec) return;  // whole waves
  const uint64_t i0 = 2 * v;
  const bool idle = skip && !((i0 >> c) & 1);
  const float4 x = idle ? make_float4(0.f, 0.f, 0.f, 0.f) : a[v];
  float4 y;
  if (t == 0) {
    y = make_float4(x.z, x.w, x.x, x.y);
  } else {
    const int d = 1 << (t - 1);
    y = make_float4(__shfl_xor(x.x, d), __shfl_xor(x.y, d), __shfl_xor(x.z, d), __shfl_xor(x.w, d));
  }
  const bool s0 = (i0 >> c) & 1, s1 = ((i0 + 1) >> c) & 1;  // control bit of each amplitude
  if (!idle) a[v] = make_float4(s0 ? y.x : x.x, s0 ? y.y : x.y, s1 ? y.z : x.z, s1 ? y.w : x.w);
}
// complex f64, t <= 5; `skip` as above
__global__ __launch_bounds__(BLOCK) void k_cnot_full_f64(double2* __restrict__ a, uint32_t c, uint32_t t,
                                                         uint64_t n, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  if (v >= n) return;  // whole waves
  const bool idle = skip && !((v >> c) & 1);
  const double2 x = idle ? make_double2(0.0, 0.0) : a[v];
  const int d = 1 << t;
  const double2 y = make_double2(shfl_xor_f64(x.x, d), shfl_xor_f64(x.y, d));
  const bool sel = (v >> c) & 1;  // per component: a select of whole vectors goes through scratch memory
  if (!idle) a[v] = make_double2(sel ? y.x : x.x, sel ? y.y : x.y);
}

// CNOT with the control inside the line and the target above the wave: lane u
// holds vectors x (bit t = 0) and y (bit t = 1) of two contiguous streams and
// exchanges the amplitudes whose control bit is set.  With `skip` (complex
// f64) lanes whose control bit is clear do nothing.
// Moves only: bit-exact.  Selects are per component (a select of whole vectors
// goes through scratch memory).
__global__ __launch_bounds__(BLOCK) void k_cnot_two_f32(float4* __restrict__ a, uint32_t c, uint32_t t,
                                                        uint64_t npairs) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= npairs) return;
  const uint64_t v0 = ins0(u, t - 1), v1 = v0 | (1ull << (t - 1));
  if (c == 0) {  // the odd amplitude of each vector
    const float4 x = a[v0], y = a[v1];
    a[v0] = make_float4(x.x, x.y, y.z, y.w);
    a[v1] = make_float4(y.x, y.y, x.z, x.w);
    return;
  }
  const bool on = (v0 >> (c - 1)) & 1;
  const float4 x = a[v0], y = a[v1];
  a[v0] = make_float4(on ? y.x : x.x, on ? y.y : x.y, on ? y.z : x.z, on ? y.w : x.w);
  a[v1] = make_float4(on ? x.x : y.x, on ? x.y : y.y, on ? x.z : y.z, on ? x.w : y.w);
}
__global__ __launch_bounds__(BLOCK) void k_cnot_two_f64(double2* __restrict__ a, uint32_t c, uint32_t t,
                                                        uint64_t npairs, bool skip) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= npairs) return;
  const uint64_t v0 = ins0(u, t), v1 = v0 | (1ull << t);
  const bool on = (v0 >> c) & 1;
  if (skip && !on) return;
  const double2 x = a[v0], y = a[v1];
  a[v0] = make_double2(on ? y.x : x.x, on ? y.y : x.y);
  a[v1] = make_double2(on ? x.x : y.x, on ? x.y : y.y);
}

// ---- CZ / CNOT over the half the gate changes -----------------------------
// When the higher CZ qubit (the CNOT control) h lies above the 128-byte line
// (complex f32: h >= 4; complex f64: h >= 3), only the lines with bit h set
// change.  These kernels walk just that half: half-space vector u maps to
// vector ins0(u, hv) | 2^hv, hv = the bit of h in the vector index, so a wave
// still reads 64 consecutive vectors in runs of >= 128 bytes.  A CNOT partner
// differs in vector bit tv (t - log2(amplitudes per vector)), which is
// half-space bit tv or tv - 1 (below / above hv): the lane shuffle uses that
// bit.  With `skip`, CZ lanes whose vector the gate leaves alone neither load
// nor store; that pays when the changed runs are >= 64 bytes (measured: at
// 16 and 32 bytes the partial writes cost more than the bytes saved).  Same
// operations as k_cz / k_cnot.  Need the half-space vector count to be a
// multiple of 64.
__global__ __launch_bounds__(BLOCK) void k_cz_half_f32(float4* __restrict__ a, uint32_t lo, uint32_t hi,
                                                       uint64_t nhalf, bool skip) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= nhalf) return;
  const uint64_t v = ins0(u, hi - 1) | (1ull << (hi - 1)), i0 = 2 * v;
  if (skip && !((i0 >> lo) & 1) && !(((i0 + 1) >> lo) & 1)) return;
  float4 x = a[v];
  if ((i0 >> lo) & 1) {
    const Cx<float> r = cz_neg(Cx<float>{x.x, x.y});
    x.x = r.re;
    x.y = r.im;
  }
  if (((i0 + 1) >> lo) & 1) {
    const Cx<float> r = cz_neg(Cx<float>{x.z, x.w});
    x.z = r.re;
    x.w = r.im;
  }
  a[v] = x;
}
__global__ __launch_bounds__(BLOCK) void k_cz_half_f64(double2* __restrict__ a, uint32_t lo, uint32_t hi,
                                                       uint64_t nhalf, bool skip) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= nhalf) return;
  const uint64_t v = ins0(u, hi) | (1ull << hi);
  const bool chg = (v >> lo) & 1;
  if (skip && !chg) return;
  const double2 x = a[v];
  const Cx<double> r = cz_neg(Cx<double>{x.x, x.y});
  a[v] = chg ? make_double2(r.re, r.im) : x;
}
// complex f32, c >= 4; lb = the half-space bit of the partner (t >= 1), <= 5
__global__ __launch_bounds__(BLOCK) void k_cnot_half_f32(float4* __restrict__ a, uint32_t c, uint32_t t,
                                                         uint32_t lb, uint64_t nhalf) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= nhalf) return;  // whole waves
  const uint64_t v = ins0(u, c - 1) | (1ull << (c - 1));
  const float4 x = a[v];
  if (t == 0) {
    a[v] = make_float4(x.z, x.w, x.x, x.y);
  } else {
    const int d = 1 << lb;
    a[v] = make_float4(__shfl_xor(x.x, d), __shfl_xor(x.y, d), __shfl_xor(x.z, d), __shfl_xor(x.w, d));
  }
}
// complex f64, c >= 3; lb <= 5 as above
__global__ __launch_bounds__(BLOCK) void k_cnot_half_f64(double2* __restrict__ a, uint32_t c, uint32_t lb,
                                                         uint64_t nhalf) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= nhalf) return;  // whole waves
  const uint64_t v = ins0(u, c) | (1ull << c);
  const double2 x = a[v];
  const int d = 1 << lb;
  a[v] = make_double2(shfl_xor_f64(x.x, d), shfl_xor_f64(x.y, d));
}

// ---- SWAP with a qubit inside the 128-byte line --------------------------
// Every line then holds amplitudes that move, so the gate is a streaming pass.
// k_swap_wave_*: both qubits inside the wave's vectors (complex f32 <= 6,
// complex f64 <= 5); a lane whose vector holds amplitudes with the two bits
// unequal takes its partner vector from lane l ^ (2^L | 2^H) (L, H the two
// bits' vector bits; the shuffle distance varies per lane).  Complex f32 with
// lo = 0: the two amplitudes of a vector differ in bit 0, so one of them moves
// to / from the other half of partner vector v ^ 2^(hi-1).
// k_swap_two_*: the high qubit above the wave; lane u holds vectors x (bit
// hi = 0) and y (bit hi = 1) of two contiguous streams.  The amplitudes that
// move are those of x with bit lo = 1 and of y with bit lo = 0: each lane sends
// its partner (lane ^ 2^L) the vector the partner needs and receives the one
// it needs, so one shuffle of one vector per lane.  With `skip` a lane stores
// only the vector that changed, and in k_swap_wave_* lanes whose vector does
// not change skip load and store (worth it for changed runs >= 64 bytes).
// Moves only: bit-exact.  Need the lane count to be a multiple of 64.
__global__ __launch_bounds__(BLOCK) void k_swap_wave_f32(float4* __restrict__ a, uint32_t lo, uint32_t hi,
                                                         uint64_t nvec, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  if (v >= nvec) return;  // whole waves
  if (lo == 0) {
    const float4 x = a[v];
    const int d = 1 << (hi - 1);
    const float4 p = make_float4(__shfl_xor(x.x, d), __shfl_xor(x.y, d), __shfl_xor(x.z, d), __shfl_xor(x.w, d));
    // bit hi = 1: amplitude 2v (bit 0 = 0) moves; bit hi = 0: amplitude 2v + 1
    const bool h = (v >> (hi - 1)) & 1;
    a[v] = make_float4(h ? p.z : x.x, h ? p.w : x.y, h ? x.z : p.x, h ? x.w : p.y);
    return;
  }
  const uint32_t bl = (v >> (lo - 1)) & 1, bh = (v >> (hi - 1)) & 1;
  const bool idle = skip && bl == bh;  // unchanged; its partner is itself
  const float4 x = idle ? make_float4(0.f, 0.f, 0.f, 0.f) : a[v];
  const int d = bl != bh ? (1 << (lo - 1)) | (1 << (hi - 1)) : 0;
  const float4 y = make_float4(__shfl_xor(x.x, d), __shfl_xor(x.y, d), __shfl_xor(x.z, d), __shfl_xor(x.w, d));
  if (!idle) a[v] = y;
}
__global__ __launch_bounds__(BLOCK) void k_swap_wave_f64(double2* __restrict__ a, uint32_t lo, uint32_t hi,
                                                         uint64_t n, bool skip) {
  const uint64_t v = flat_block() * BLOCK + threadIdx.x;
  if (v >= n) return;  // whole waves
  const uint32_t bl = (v >> lo) & 1, bh = (v >> hi) & 1;
  const bool idle = skip && bl == bh;
  const double2 x = idle ? make_double2(0.0, 0.0) : a[v];
  const int d = bl != bh ? (1 << lo) | (1 << hi) : 0;
  const double2 y = make_double2(shfl_xor_f64(x.x, d), shfl_xor_f64(x.y, d));
  if (!idle) a[v] = y;
}
// complex f32, lo <= 6, hi >= 7
__global__ __launch_bounds__(BLOCK) void k_swap_two_f32(float4* __restrict__ a, uint32_t lo, uint32_t hi,
                                                        uint64_t npairs, bool skip) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= npairs) return;  // whole waves
  const uint64_t v0 = ins0(u, hi - 1), v1 = v0 | (1ull << (hi - 1));
  const float4 x = a[v0], y = a[v1];
  if (lo == 0) {  // x.hi (bit 0 = 1, bit hi = 0) <-> y.lo (bit 0 = 0, bit hi = 1)
    a[v0] = make_float4(x.x, x.y, y.x, y.y);
    a[v1] = make_float4(x.z, x.w, y.z, y.w);
    return;
  }
  const int d = 1 << (lo - 1);
  const bool one = (u >> (lo - 1)) & 1;  // bit lo of both vectors
  // what the partner lane needs (per component: a select of whole vectors goes through scratch memory)
  const float4 s = make_float4(one ? x.x : y.x, one ? x.y : y.y, one ? x.z : y.z, one ? x.w : y.w);
  const float4 r = make_float4(__shfl_xor(s.x, d), __shfl_xor(s.y, d), __shfl_xor(s.z, d), __shfl_xor(s.w, d));
  if (one) {
    a[v0] = r;  // x's amplitudes with bit lo = 1 take y's with bit lo = 0
    if (!skip) a[v1] = y;
  } else {
    if (!skip) a[v0] = x;
    a[v1] = r;
  }
}
// complex f64, lo <= 5, hi >= 6
__global__ __launch_bounds__(BLOCK) void k_swap_two_f64(double2* __restrict__ a, uint32_t lo, uint32_t hi,
                                                        uint64_t npairs, bool skip) {
  const uint64_t u = flat_block() * BLOCK + threadIdx.x;
  if (u >= npairs) return;  // whole waves
  const uint64_t v0 = ins0(u, hi), v1 = v0 | (1ull << hi);
  const double2 x = a[v0], y = a[v1];
  const int d = 1 << lo;
  const bool one = (u >> lo) & 1;
  const double2 s = make_double2(one ? x.x : y.x, one ? x.y : y.y);
  const double2 r = make_double2(shfl_xor_f64(s.x, d), shfl_xor_f64(s.y, d));
  if (one) {
    a[v0] = r;
    if (!skip) a[v1] = y;
  } else {
    if (!skip) a[v0] = x;
    a[v1] = r;
  }
}

// ---- diagonal 2q (Rzz): a_i *= d[b1*2 + b2], f32 float4 / f64 double2 ----
__global__ __launch_bounds__(BLOCK) void k_diag2_f32(float4* __restrict__ a, uint32_t q1, uint32_t q2,
                                                     uint64_t nvec, M4<float> d) {
  uint64_t g = flat_block() * BLOCK + threadIdx.x;
  if (g >= nvec) return;
  float4 x = a[g];
  uint64_t i0 = 2 * g, i1 = 2 * g + 1;
  int k0 = (int)(((i0 >> q1) & 1) * 2 + ((i0 >> q2) & 1));
  int k1 = (int)(((i1 >> q1) & 1) * 2 + ((i1 >> q2) & 1));
  Cx<float> r0 = cmul(d.m[k0 * 5], Cx<float>{x.x, x.y});
  Cx<float> r1 = cmul(d.m[k1 * 5], Cx<float>{x.z, x.w});
  a[g] = make_float4(r0.re, r0.im, r1.re, r1.im);
}

__global__ __launch_bounds__(BLOCK) void k_diag2_f64(double2* __restrict__ a, uint32_t q1, uint32_t q2, uint64_t n,
                                                     M4<double> d) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  double2 x = a[i];
  int k = (int)(((i >> q1) & 1) * 2 + ((i >> q2) & 1));
  Cx<double> r = cmul(d.m[k * 5], Cx<double>{x.x, x.y});
  a[i] = make_double2(r.re, r.im);
}

// ---- |0..0> -----------------------------------------------------------------
template <class V>
__global__ __launch_bounds__(BLOCK) void k_init_zero(V* __restrict__ a, uint64_t n) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  V z{};
  if (i == 0) z.x = 1;
  a[i] = z;
}

// ---- measurement collapse (state.rs:310-317, mixed_precision.rs:499-506) ---
template <class T, class V>
__global__ __launch_bounds__(BLOCK) void k_collapse(V* __restrict__ a, uint32_t q, int result, T nf, uint64_t n) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  V x = a[i];
  int one = (int)((i >> q) & 1);
  if (one == result) {
    x.x = x.x * nf;
    x.y = x.y * nf;
  } else {
    x.x = 0;
    x.y = 0;
  }
  a[i] = x;
}

// ---- scale every amplitude (normalize, state.rs:419-421) -------------------
template <class T, class V>
__global__ __launch_bounds__(BLOCK) void k_scale(V* __restrict__ a, T f, uint64_t n) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  V x = a[i];
  x.x = x.x * f;
  x.y = x.y * f;
  a[i] = x;
}

// ---- sign flips (grover.rs:118-129 oracle, :136-143 diffusion) ------------
// Negates a[i] for i in [start, n): a sign flip of both parts, exact.
template <class V>
__global__ __launch_bounds__(BLOCK) void k_negate_from(V* __restrict__ a, uint64_t start, uint64_t n) {
  uint64_t i = start + flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  V x = a[i];
  x.x = -x.x;
  x.y = -x.y;
  a[i] = x;
}
// Negates a[idx[k]] for k < count; the indices are distinct (the host folds
// repeated indices: an index listed twice is negated twice, i.e. not at all).
template <class V>
__global__ __launch_bounds__(BLOCK) void k_negate_listed(V* __restrict__ a, const uint64_t* __restrict__ idx,
                                                         uint64_t count) {
  uint64_t k = flat_block() * BLOCK + threadIdx.x;
  if (k >= count) return;
  V x = a[idx[k]];
  x.x = -x.x;
  x.y = -x.y;
  a[idx[k]] = x;
}

// ---- plain 16 B copy (bandwidth reference and state copies) ----------------
template <int U, bool NT>
__global__ __launch_bounds__(BLOCK) void k_copy16(float4* __restrict__ dst, const float4* __restrict__ src,
                                                  uint64_t nvec) {
  const uint64_t g0 = flat_block() * U * BLOCK + threadIdx.x;
  float4 x[U];
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < nvec) x[k] = ldg<NT>(src + g);
  }
#pragma unroll
  for (int k = 0; k < U; ++k) {
    uint64_t g = g0 + (uint64_t)k * BLOCK;
    if (g < nvec) stg<NT>(dst + g, x[k]);
  }
}

inline dim3 grid_for(uint64_t items, uint64_t per_block) { return grid_blocks((items + per_block - 1) / per_block); }

template <class T>
M2<T> to_m2(const rq_c64* m) {
  M2<T> r;
  r.m00 = Cx<T>{(T)m[0].re, (T)m[0].im};
  r.m01 = Cx<T>{(T)m[1].re, (T)m[1].im};
  r.m10 = Cx<T>{(T)m[2].re, (T)m[2].im};
  r.m11 = Cx<T>{(T)m[3].re, (T)m[3].im};
  return r;
}
template <class T>
M4<T> to_m4(const rq_c64* m) {
  M4<T> r;
  for (int k = 0; k < 16; ++k) r.m[k] = Cx<T>{(T)m[k].re, (T)m[k].im};
  return r;
}

// Block chunk of the two-stream 1q kernels by log2 of the distance between the
// streams in bytes (2^q * amplitude size), measured on MI355X with a 30-qubit
// complex-f32 state (non-temporal access, one group per thread, partner loads
// issued after the first ones returned) over chunks 0 (linear order) and
// 32 .. 2048: 32 blocks (128 KB of each stream per XCD at a time) is best up
// to 256 KB apart; from 8 MB on, 32 and 64 alternate with period 4 in log2 of
// the distance (32 loses 0.6 TB/s at 8 MB, 128 MB and 2 GB, 64 as much at
// 16 MB, 256 MB and 4 GB).
constexpr uint16_t XCD_CHUNK_BY_LOG2_DIST[64] = {
    32,  32,  32,  32,  32,  32,  32,  32,  32,  32,     // 0..9
    32,  32,  32,  32,  32,  32,  32,  32,  32,  128,    // 10..19: 512 KB 128
    32,  32,  32,  64,  32,  32,  64,  64,  32,  64,     // 20..29
    64,  64,  32,  64,  64,  64,  32,  64,  64,  64,     // 30..39
    32,  64,  64,  64,  32,  64,  64,  64,  32,  64, 64, 64, 32, 64,
    64,  64,  32,  64,  64,  64,  32,  64,  64,  64};

// The chunk for a two-stream launch of nb workgroups whose streams are
// 2^log2d bytes apart: the tuned value or the table, and 0 when nb is not a
// multiple of 8 chunks.
uint32_t xcd_chunk_for(uint64_t nb, int log2d) {
  const int t = stream_tune().xcd_chunk;
  const uint64_t c = t >= 0 ? (uint64_t)t : XCD_CHUNK_BY_LOG2_DIST[log2d < 63 ? log2d : 63];
  return (c && nb % (8 * c) == 0) ? (uint32_t)c : 0u;
}
uint64_t grid_total(dim3 g) { return (uint64_t)g.x * g.y; }

template <int U, bool NT>
void launch_1q_t(State& s, uint32_t q, const rq_c64* m) {
  if (s.prec == RQ_C64 && q <= 5 && s.n >= 64) {
    hipLaunchKernelGGL((k_1q_f64_lx<U, NT>), grid_for(s.n, (uint64_t)BLOCK * U), dim3(BLOCK), 0, s.stream,
                       (double2*)s.amps, q, (uint64_t)s.n, to_m2<double>(m));
  } else if (s.prec == RQ_C64) {
    uint64_t np = s.n / 2;
    const dim3 g = grid_for(np, (uint64_t)BLOCK * U);
    hipLaunchKernelGGL((k_1q_f64<U, NT>), g, dim3(BLOCK), 0, s.stream, (double2*)s.amps, q, np, to_m2<double>(m),
                       xcd_chunk_for(grid_total(g), (int)q + 4));
  } else if (q == 0) {
    uint64_t np = s.n / 2;
    hipLaunchKernelGGL((k_1q_f32_q0<U, NT>), grid_for(np, (uint64_t)BLOCK * U), dim3(BLOCK), 0, s.stream,
                       (float4*)s.amps, np, to_m2<float>(m));
  } else if (q <= 6 && s.n >= 128) {
    uint64_t nv = s.n / 2;
    hipLaunchKernelGGL((k_1q_f32_lx<U, NT>), grid_for(nv, (uint64_t)BLOCK * U), dim3(BLOCK), 0, s.stream,
                       (float4*)s.amps, q, nv, to_m2<float>(m));
  } else {
    uint64_t ng = s.n / 4;
    const dim3 g = grid_for(ng, (uint64_t)BLOCK * U);
    hipLaunchKernelGGL((k_1q_f32_v2<U, NT>), g, dim3(BLOCK), 0, s.stream, (float4*)s.amps, q, ng, to_m2<float>(m),
                       xcd_chunk_for(grid_total(g), (int)q + 3));
  }
}

template <int U, bool NT>
void launch_copy_t(void* dst, const void* src, size_t bytes, hipStream_t st) {
  uint64_t nv = bytes / 16;
  hipLaunchKernelGGL((k_copy16<U, NT>), grid_for(nv, (uint64_t)BLOCK * U), dim3(BLOCK), 0, st, (float4*)dst,
                     (const float4*)src, nv);
}

// QuantumStateF32::from_f64 / to_f64 (mixed_precision.rs:260-288): each
// amplitude narrowed with round-to-nearest (Complex32::from_f64, `as f32`) or
// widened exactly (Complex32::to_f64).
__global__ __launch_bounds__(BLOCK) void k_narrow(float2* __restrict__ dst, const double2* __restrict__ src,
                                                  uint64_t n) {
  const uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) {
    const double2 v = src[i];
    dst[i] = make_float2((float)v.x, (float)v.y);
  }
}
__global__ __launch_bounds__(BLOCK) void k_widen(double2* __restrict__ dst, const float2* __restrict__ src,
                                                 uint64_t n) {
  const uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) {
    const float2 v = src[i];
    dst[i] = make_double2((double)v.x, (double)v.y);
  }
}

}  // namespace

// Shape of the streaming kernels: amplitude groups per thread (1, 2, 4, 8)
// and non-temporal loads/stores.  Chosen by measurement (rq_tune_set).
StreamTune& stream_tune() {
  static StreamTune t;
  return t;
}

void launch_init_zero(State& s) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_init_zero<double2>, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, s.n);
  else
    hipLaunchKernelGGL(k_init_zero<float2>, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps, s.n);
}

void launch_apply_1q(State& s, uint32_t q, const rq_c64* m) {
  const StreamTune& t = stream_tune();
  switch (t.unroll * 2 + (t.nontemporal ? 1 : 0)) {
    case 2: launch_1q_t<1, false>(s, q, m); break;
    case 3: launch_1q_t<1, true>(s, q, m); break;
    case 4: launch_1q_t<2, false>(s, q, m); break;
    case 5: launch_1q_t<2, true>(s, q, m); break;
    case 9: launch_1q_t<4, true>(s, q, m); break;
    case 16: launch_1q_t<8, false>(s, q, m); break;
    case 17: launch_1q_t<8, true>(s, q, m); break;
    default: launch_1q_t<4, false>(s, q, m); break;
  }
}

void launch_apply_2q(State& s, uint32_t q1, uint32_t q2, const rq_c64* m) {
  uint64_t nq4 = s.n / 4;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_2q_dense<double, double2>), grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream,
                       (double2*)s.amps, q1, q2, nq4, to_m4<double>(m));
  else
    hipLaunchKernelGGL((k_2q_dense<float, float2>), grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream,
                       (float2*)s.amps, q1, q2, nq4, to_m4<float>(m));
}

void launch_cnot(State& s, uint32_t c, uint32_t t) {
  uint64_t nq4 = s.n / 4;
  // a low control qubit touches every line: one streaming pass
  // bytes per run of amplitudes with the control set: from 64 up, lanes with
  // the control clear skip their load and store
  const bool skip = ((uint64_t)(s.prec == RQ_C32 ? 8 : 16) << c) >= 64;
  if (s.prec == RQ_C32 && c <= 3 && t <= 6 && s.n >= 128) {
    hipLaunchKernelGGL(k_cnot_full_f32, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, c, t,
                       (uint64_t)(s.n / 2), skip);
    return;
  }
  if (s.prec == RQ_C64 && c <= 2 && t <= 5 && s.n >= 64) {
    hipLaunchKernelGGL(k_cnot_full_f64, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, c, t,
                       (uint64_t)s.n, skip);
    return;
  }
  // t >= 7 (control 3: k_cnot measured faster, 2.78-3.00 against 2.83-3.05 ms)
  if (s.prec == RQ_C32 && c <= 2 && s.n >= 256) {
    hipLaunchKernelGGL(k_cnot_two_f32, grid_for(s.n / 4, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, c, t,
                       (uint64_t)(s.n / 4));
    return;
  }
  if (s.prec == RQ_C64 && c <= 2 && s.n >= 128) {  // t >= 6
    hipLaunchKernelGGL(k_cnot_two_f64, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, c, t,
                       (uint64_t)(s.n / 2), skip);
    return;
  }
  // a control above the line and a target within the wave: the control half
  // only (complex f32: k_cnot measured as fast or faster from control 6 up)
  if (s.prec == RQ_C32 && c >= 4 && c <= 5 && s.n >= 256) {
    const uint32_t lb = t == 0 ? 0 : (t < c ? t - 1 : t - 2);
    if (lb <= 5) {
      hipLaunchKernelGGL(k_cnot_half_f32, grid_for(s.n / 4, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, c,
                         t, lb, (uint64_t)(s.n / 4));
      return;
    }
  }
  if (s.prec == RQ_C64 && c >= 3 && s.n >= 128) {
    const uint32_t lb = t < c ? t : t - 1;
    if (lb <= 5) {
      hipLaunchKernelGGL(k_cnot_half_f64, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, c,
                         lb, (uint64_t)(s.n / 2));
      return;
    }
  }
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_cnot<double2>, grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, c, t, nq4);
  else
    hipLaunchKernelGGL(k_cnot<float2>, grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps, c, t, nq4);
}

void launch_swap(State& s, uint32_t q1, uint32_t q2) {
  uint64_t nq4 = s.n / 4;
  const uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  const uint64_t run = (uint64_t)(s.prec == RQ_C32 ? 8 : 16) << lo;  // bytes per run of moved amplitudes
  // a qubit inside the 128-byte line: every line changes, one streaming pass
  if (s.prec == RQ_C32 && lo <= 3 && s.n >= 256) {
    if (hi <= 6)
      hipLaunchKernelGGL(k_swap_wave_f32, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, lo, hi,
                         (uint64_t)(s.n / 2), run >= 64);
    else
      hipLaunchKernelGGL(k_swap_two_f32, grid_for(s.n / 4, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, lo, hi,
                         (uint64_t)(s.n / 4), run >= 64);
    return;
  }
  if (s.prec == RQ_C64 && lo <= 2 && s.n >= 128) {
    if (hi <= 5)
      hipLaunchKernelGGL(k_swap_wave_f64, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, lo, hi,
                         (uint64_t)s.n, run >= 64);
    else
      hipLaunchKernelGGL(k_swap_two_f64, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, lo,
                         hi, (uint64_t)(s.n / 2), run >= 64);
    return;
  }
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_swap<double2>, grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, q1, q2, nq4);
  else
    hipLaunchKernelGGL(k_swap<float2>, grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps, q1, q2, nq4);
}

void launch_cz(State& s, uint32_t q1, uint32_t q2) {
  uint64_t nq4 = s.n / 4;
  const uint32_t lo = q1 < q2 ? q1 : q2, hi = q1 < q2 ? q2 : q1;
  // bytes in one run of changed amplitudes (2^lo amplitudes)
  const uint64_t run = (uint64_t)(s.prec == RQ_C32 ? 8 : 16) << lo;
  // a low qubit and one above the line: the half with the high bit set
  if (s.prec == RQ_C32 && lo <= 3 && hi >= 4 && s.n >= 256) {
    hipLaunchKernelGGL(k_cz_half_f32, grid_for(s.n / 4, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, lo, hi,
                       (uint64_t)(s.n / 4), run >= 64);
    return;
  }
  if (s.prec == RQ_C64 && lo <= 2 && hi >= 3 && s.n >= 128) {
    hipLaunchKernelGGL(k_cz_half_f64, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, lo, hi,
                       (uint64_t)(s.n / 2), run >= 64);
    return;
  }
  if (s.prec == RQ_C32 && lo <= 3 && s.n >= 128) {
    hipLaunchKernelGGL(k_cz_full_f32, grid_for(s.n / 2, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, q1, q2,
                       (uint64_t)(s.n / 2), run >= 32);
    return;
  }
  if (s.prec == RQ_C64 && lo <= 2 && s.n >= 64) {
    hipLaunchKernelGGL(k_cz_full_f64, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, q1, q2,
                       (uint64_t)s.n, run >= 32);
    return;
  }
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_cz<double, double2>), grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, q1,
                       q2, nq4);
  else
    hipLaunchKernelGGL((k_cz<float, float2>), grid_for(nq4, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps, q1, q2,
                       nq4);
}

void launch_diag_2q(State& s, uint32_t q1, uint32_t q2, const rq_c64* d4) {
  if (s.prec == RQ_C64) {
    hipLaunchKernelGGL(k_diag2_f64, grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, q1, q2, s.n,
                       to_m4<double>(d4));
  } else {
    uint64_t nv = s.n / 2;
    hipLaunchKernelGGL(k_diag2_f32, grid_for(nv, BLOCK), dim3(BLOCK), 0, s.stream, (float4*)s.amps, q1, q2, nv,
                       to_m4<float>(d4));
  }
}

void launch_collapse(State& s, uint32_t q, int result, double nf64) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_collapse<double, double2>), grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream,
                       (double2*)s.amps, q, result, nf64, s.n);
  else
    hipLaunchKernelGGL((k_collapse<float, float2>), grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps,
                       q, result, (float)nf64, s.n);
}

void launch_scale(State& s, double f) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_scale<double, double2>), grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps,
                       f, s.n);
  else
    hipLaunchKernelGGL((k_scale<float, float2>), grid_for(s.n, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps,
                       (float)f, s.n);
}

void launch_negate_from(State& s, uint64_t start) {
  if (start >= s.n) return;
  const uint64_t m = s.n - start;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_negate_from<double2>, grid_for(m, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps, start,
                       (uint64_t)s.n);
  else
    hipLaunchKernelGGL(k_negate_from<float2>, grid_for(m, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps, start,
                       (uint64_t)s.n);
}

void launch_negate_listed(State& s, const uint64_t* d_idx, uint64_t count) {
  if (!count) return;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_negate_listed<double2>, grid_for(count, BLOCK), dim3(BLOCK), 0, s.stream, (double2*)s.amps,
                       d_idx, count);
  else
    hipLaunchKernelGGL(k_negate_listed<float2>, grid_for(count, BLOCK), dim3(BLOCK), 0, s.stream, (float2*)s.amps,
                       d_idx, count);
}

void launch_plain_copy(void* dst, const void* src, size_t bytes, hipStream_t st) {
  const StreamTune& t = stream_tune();
  switch (t.unroll * 2 + (t.nontemporal ? 1 : 0)) {
    case 2: launch_copy_t<1, false>(dst, src, bytes, st); break;
    case 3: launch_copy_t<1, true>(dst, src, bytes, st); break;
    case 4: launch_copy_t<2, false>(dst, src, bytes, st); break;
    case 5: launch_copy_t<2, true>(dst, src, bytes, st); break;
    case 9: launch_copy_t<4, true>(dst, src, bytes, st); break;
    case 16: launch_copy_t<8, false>(dst, src, bytes, st); break;
    case 17: launch_copy_t<8, true>(dst, src, bytes, st); break;
    default: launch_copy_t<4, false>(dst, src, bytes, st); break;
  }
}

void launch_copy(State& dst, const State& src) { launch_plain_copy(dst.amps, src.amps, src.bytes(), dst.stream); }

void launch_convert(State& dst, const State& src) {
  if (dst.prec == RQ_C32)
    hipLaunchKernelGGL(k_narrow, grid_for(src.n, BLOCK), dim3(BLOCK), 0, dst.stream, (float2*)dst.amps,
                       (const double2*)src.amps, (uint64_t)src.n);
  else
    hipLaunchKernelGGL(k_widen, grid_for(src.n, BLOCK), dim3(BLOCK), 0, dst.stream, (double2*)dst.amps,
                       (const float2*)src.amps, (uint64_t)src.n);
}

}  // namespace rq
