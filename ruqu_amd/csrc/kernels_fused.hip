// Fused stage kernel: one workgroup per tile of 2^m amplitudes, every gate of
// the stage applied in registers, one HBM read and one HBM write per
// amplitude per stage (layout and planning in fusion.h / fusion.cpp).
//
// Thread t of a phase holds the 16 amplitudes whose tile index is
//   tile(k, t) = sum_b bit_b(k) << regt[b]  |  sum_j bit_j(t) << thrt[j],
// k = 0..15 over the 4 register qubits.  Phases change the register qubits
// by writing the tile to LDS and reading it back with the next mapping.
// A stage's first phase reads HBM and its last phase writes it directly in
// their own mappings (phase.direct, set whenever a workgroup holds full
// waves): the lanes hold the lowest non-register tile bits, so a thread's 16
// accesses together cover every line they touch.  Otherwise the tile passes
// through the natural mapping (thread bits = the low tile bits) and LDS.
//
// Arithmetic: FP contraction is allowed here (fused mode), so results differ
// from the per-gate kernels at the rounding level.
#include <array>
#include <cstddef>
#include <cstring>
#include <utility>

#include "fusion.h"

#pragma clang fp contract(fast)

namespace rq {
namespace {

template <class T>
struct alignas(2 * sizeof(T)) Cx {  // one 8- / 16-byte LDS access per amplitude
  T re, im;
};

template <class T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> m, Cx<T> a) {
  return Cx<T>{m.re * a.re - m.im * a.im, m.re * a.im + m.im * a.re};
}
template <class T>
__device__ __forceinline__ Cx<T> cmadd(Cx<T> m0, Cx<T> a0, Cx<T> m1, Cx<T> a1) {
  return Cx<T>{m0.re * a0.re - m0.im * a0.im + m1.re * a1.re - m1.im * a1.im,
               m0.re * a0.im + m0.im * a0.re + m1.re * a1.im + m1.im * a1.re};
}

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// LDS byte address of tile index t.  The XOR swizzle folds index bits 5..9
// (f32; 4..7 for f64) into the bank bits, so the lanes of a wave hit distinct
// banks whichever tile bits they hold.  It is linear over bits: the address of
// (thread part | register part) is the XOR of the two parts' addresses, so a
// phase computes its per-thread part once and each register adds one XOR.
template <class T>
__host__ __device__ __forceinline__ uint32_t lds_addr(uint32_t t) {
  if constexpr (sizeof(T) == 4) return (t ^ ((t >> 5) & 31u)) * 8u;
  else return (t ^ ((t >> 4) & 15u)) * 16u;
}

// Per-thread part of the LDS address, hidden from the optimizer: otherwise it
// factors the common *8 out of (thread part ^ register part) and each access
// pays a shift on top of the XOR.
template <class T>
__device__ __forceinline__ uint32_t lane_lds_addr(uint32_t t) {
  uint32_t a = lds_addr<T>(t);
  asm("" : "+v"(a));
  return a;
}

// Every op reads the 16 amplitudes from x and writes all 16 to y, a distinct
// register set.  apply_ops alternates the two sets (x -> y, then y -> x), so
// an op's results land in fresh registers and no copy back is needed between
// ops; entries an op leaves unchanged cost one move.
template <class T>
using Amps = Cx<T>[FUSE_R];

// a with the sign bits of both parts XORed with `sign` (0 or 0x80000000).
template <class T>
__device__ __forceinline__ Cx<T> flip(Cx<T> a, uint32_t sign) {
  if constexpr (sizeof(T) == 4) {
    return Cx<T>{__uint_as_float(__float_as_uint(a.re) ^ sign), __uint_as_float(__float_as_uint(a.im) ^ sign)};
  } else {
    const unsigned long long s = (unsigned long long)sign << 32;
    return Cx<T>{__longlong_as_double(__double_as_longlong(a.re) ^ s),
                 __longlong_as_double(__double_as_longlong(a.im) ^ s)};
  }
}

// A copy of a in new registers.  Ops that pass an amplitude through unchanged
// (permutations, partial sign flips, the padding record) copy it explicitly:
// if an op's output were the input value itself, the register allocator would
// merge the two register sets at the join of the op switch and then have to
// copy the results of every other op instead.
__device__ __forceinline__ uint32_t fresh32(uint32_t a) {
  uint32_t r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(a));
  return r;
}
__device__ __forceinline__ uint64_t fresh64(uint64_t a) {
  uint64_t r;
  asm("v_mov_b64 %0, %1" : "=v"(r) : "v"(a));
  return r;
}
template <class T>
__device__ __forceinline__ Cx<T> fresh(Cx<T> a) {
  if constexpr (sizeof(T) == 4) {
    const uint64_t u = ((uint64_t)__float_as_uint(a.im) << 32) | __float_as_uint(a.re);
    const uint64_t r = fresh64(u);
    return Cx<T>{__uint_as_float((uint32_t)r), __uint_as_float((uint32_t)(r >> 32))};
  } else {
    return Cx<T>{__longlong_as_double(fresh64(__double_as_longlong(a.re))),
                 __longlong_as_double(fresh64(__double_as_longlong(a.im)))};
  }
}

// 2x2 on register bit B.
template <int B, class T>
__device__ __forceinline__ void u1(const Amps<T>& x, Amps<T>& y, const Cx<T>* __restrict__ mp) {
  const Cx<T> m00 = mp[0], m01 = mp[1], m10 = mp[2], m11 = mp[3];
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    y[k] = cmadd(m00, x[k], m01, x[k1]);
    y[k1] = cmadd(m10, x[k], m11, x[k1]);
  }
}

// 4x4 on register bits (B0 = q1, row-high bit; B1 = q2, row-low bit).
template <class T>
__device__ __forceinline__ Cx<T> row4(const Cx<T>* __restrict__ mp, int r, Cx<T> x0, Cx<T> x1, Cx<T> x2, Cx<T> x3) {
  Cx<T> s = cmadd(mp[r * 4 + 0], x0, mp[r * 4 + 1], x1);
  Cx<T> s2 = cmadd(mp[r * 4 + 2], x2, mp[r * 4 + 3], x3);
  return Cx<T>{s.re + s2.re, s.im + s2.im};
}
template <int B0, int B1, class T>
__device__ __forceinline__ void u2(const Amps<T>& x, Amps<T>& y, const Cx<T>* __restrict__ mp) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & ((1 << B0) | (1 << B1))) continue;
    const int i0 = k, i1 = k | (1 << B1), i2 = k | (1 << B0), i3 = k | (1 << B0) | (1 << B1);
    y[i0] = row4(mp, 0, x[i0], x[i1], x[i2], x[i3]);
    y[i1] = row4(mp, 1, x[i0], x[i1], x[i2], x[i3]);
    y[i2] = row4(mp, 2, x[i0], x[i1], x[i2], x[i3]);
    y[i3] = row4(mp, 3, x[i0], x[i1], x[i2], x[i3]);
  }
}

// CNOT target on register bit B: with register control bit C, unconditional,
// or under a per-thread / per-tile condition c.
template <int B, int C, class T>
__device__ __forceinline__ void cx_rc(const Amps<T>& x, Amps<T>& y) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = fresh(((k >> C) & 1) ? x[k ^ (1 << B)] : x[k]);
}
template <int B, class T>
__device__ __forceinline__ void cx_all(const Amps<T>& x, Amps<T>& y) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = fresh(x[k ^ (1 << B)]);
}
// Per lane: b where the lane's bit of `lanes` is set, else a.  One
// v_cndmask per 32-bit word: a plain conditional between two entries of the
// register array would be folded into an indexed access to it, which puts the
// array in scratch memory.
__device__ __forceinline__ uint32_t pick32(uint32_t a, uint32_t b, uint64_t lanes) {
  uint32_t r;
  asm("v_cndmask_b32 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(lanes));
  return r;
}
__device__ __forceinline__ float pick(float a, float b, uint64_t lanes) {
  return __uint_as_float(pick32(__float_as_uint(a), __float_as_uint(b), lanes));
}
__device__ __forceinline__ double pick(double a, double b, uint64_t lanes) {
  const unsigned long long ua = __double_as_longlong(a), ub = __double_as_longlong(b);
  const uint32_t lo = pick32((uint32_t)ua, (uint32_t)ub, lanes), hi = pick32((uint32_t)(ua >> 32), (uint32_t)(ub >> 32), lanes);
  return __longlong_as_double(((unsigned long long)hi << 32) | lo);
}
template <int B, class T>
__device__ __forceinline__ void cx_sel(const Amps<T>& x, Amps<T>& y, uint64_t lanes) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    const Cx<T> a = x[k], b = x[k ^ (1 << B)];
    y[k] = Cx<T>{pick(a.re, b.re, lanes), pick(a.im, b.im, lanes)};
  }
}

// The same permutations after a relative phase rho on the amplitudes whose
// register bit B is set (the pending factor of the target bit moved in front
// of the CNOT, so that it leaves no coupling behind).
template <int B, int C, class T>
__device__ __forceinline__ void cx_rc_ph(const Amps<T>& x, Amps<T>& y, Cx<T> rho) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    const int src = ((k >> C) & 1) ? k ^ (1 << B) : k;
    y[k] = ((src >> B) & 1) ? cmul(rho, x[src]) : fresh(x[src]);
  }
}
template <int B, class T>
__device__ __forceinline__ void cx_sel_ph(const Amps<T>& x, Amps<T>& y, uint64_t lanes, Cx<T> rho) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    const Cx<T> a = x[k], b = cmul(rho, x[k1]);
    y[k] = Cx<T>{pick(a.re, b.re, lanes), pick(a.im, b.im, lanes)};
    y[k1] = Cx<T>{pick(b.re, a.re, lanes), pick(b.im, a.im, lanes)};
  }
}

// CZ: negate where register bits P0 and P1 are set; where register bit P is
// set (sign applied: the other bit's value << 31); everywhere (sign).
template <int P0, int P1, class T>
__device__ __forceinline__ void cz_rr(const Amps<T>& x, Amps<T>& y) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = (((k >> P0) & (k >> P1)) & 1) ? flip(x[k], 0x80000000u) : fresh(x[k]);
}
template <int P, class T>
__device__ __forceinline__ void cz_r(const Amps<T>& x, Amps<T>& y, uint32_t sign) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = ((k >> P) & 1) ? flip(x[k], sign) : fresh(x[k]);
}
template <class T>
__device__ __forceinline__ void flip_all(const Amps<T>& x, Amps<T>& y, uint32_t sign) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = flip(x[k], sign);
}

// Diagonal factors: selected by register bit P; indexed by register bits
// (P0, P1); one factor for all 16 amplitudes.
template <int P, class T>
__device__ __forceinline__ void dsel_r(const Amps<T>& x, Amps<T>& y, Cx<T> f0, Cx<T> f1) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = cmul(((k >> P) & 1) ? f1 : f0, x[k]);
}
template <int P0, int P1, class T>
__device__ __forceinline__ void d2_rr(const Amps<T>& x, Amps<T>& y, const Cx<T>* __restrict__ mp) {
  const Cx<T> f0 = mp[0], f1 = mp[1], f2 = mp[2], f3 = mp[3];
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    const int i = ((k >> P0) & 1) * 2 + ((k >> P1) & 1);
    y[k] = cmul(i == 0 ? f0 : (i == 1 ? f1 : (i == 2 ? f2 : f3)), x[k]);
  }
}
template <class T>
__device__ __forceinline__ void dall(const Amps<T>& x, Amps<T>& y, Cx<T> f) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = cmul(f, x[k]);
}
template <class T>
__device__ __forceinline__ void dtab(const Amps<T>& x, Amps<T>& y, const Cx<T>* __restrict__ mp) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) y[k] = cmul(mp[k], x[k]);
}

// a + c b with one complex coefficient: 4 FMAs.
template <class T>
__device__ __forceinline__ Cx<T> axpy1(Cx<T> a, T cr, T ci, Cx<T> b) {
  return Cx<T>{__builtin_fma(-ci, b.im, __builtin_fma(cr, b.re, a.re)),
               __builtin_fma(ci, b.re, __builtin_fma(cr, b.im, a.im))};
}
template <>
__device__ __forceinline__ Cx<float> axpy1(Cx<float> a, float cr, float ci, Cx<float> b) {
  return Cx<float>{__builtin_fmaf(-ci, b.im, __builtin_fmaf(cr, b.re, a.re)),
                   __builtin_fmaf(ci, b.re, __builtin_fmaf(cr, b.im, a.im))};
}

// Factored 2x2 on register bit B: [[1, r], [s, 1]] (8 FMAs per pair instead
// of 16) or, with ANTI, [[r, 1], [1, s]].  Every output is a + c b in two
// FMAs; all first FMAs are issued before any second one.  Each input then
// stays live until a second FMA reads it as a multiplicand, so no first FMA
// can overwrite its addend's register (two-address form): the outputs are
// written straight into the other register set, with no copies.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
template <class T>
__device__ __forceinline__ void axpy_first(Cx<T>& t, Cx<T> a, T cr, Cx<T> b) {
  t.re = fma_t(cr, b.re, a.re);
  t.im = fma_t(cr, b.im, a.im);
}
template <class T>
__device__ __forceinline__ void axpy_second(Cx<T>& t, T ci, Cx<T> b) {
  t.re = fma_t(-ci, b.im, t.re);
  t.im = fma_t(ci, b.re, t.im);
}
template <int B, bool ANTI, class T>
__device__ __forceinline__ void u1f_t(const Amps<T>& x, Amps<T>& y, T rr, T ri, T sr, T si) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    axpy_first(y[k], ANTI ? x[k1] : x[k], rr, ANTI ? x[k] : x[k1]);
    axpy_first(y[k1], ANTI ? x[k] : x[k1], sr, ANTI ? x[k1] : x[k]);
  }
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    axpy_second(y[k], ri, ANTI ? x[k] : x[k1]);
    axpy_second(y[k1], si, ANTI ? x[k1] : x[k]);
  }
}
template <int B, class T>
__device__ __forceinline__ void u1f(const Amps<T>& x, Amps<T>& y, T rr, T ri, T sr, T si) {
  u1f_t<B, false>(x, y, rr, ri, sr, si);
}
template <int B, class T>
__device__ __forceinline__ void u1f_anti(const Amps<T>& x, Amps<T>& y, T rr, T ri, T sr, T si) {
  u1f_t<B, true>(x, y, rr, ri, sr, si);
}

// Packed op record (built from FOp at launch; two records are one 64-byte
// scalar load):
//   x: kind:3 | t0:2 | t1:2 | cloc:2 | cpos:5 | dloc0:2 | dpos0:6 | dloc1:2 | dpos1:6
//   y: matrix / diagonal offset in the pool
//   c: for a factored 2x2 on a complex-f32 state, its coefficients r and s
//      (filled in at upload), so the op needs no second, dependent load
//   tmask: for a factored 2x2, its form (bit 31: unit anti-diagonal) and the
//      tile-id bit that toggles it; the form of a tile is the parity of
//      tmask & (blockIdx.x | 1 << 31)
// Locations: 0 none, LOC_REG + 1 register bit, LOC_THR + 1 thread bit,
// LOC_OUT + 1 tile-id bit (position = bit of blockIdx.x).
struct POp {
  uint32_t x, y;
  float c[4];
  uint32_t tmask;  // factored 2x2: form bit 31, tile-id bit of the form toggle (0: none)
  uint32_t pad;
};
static_assert(sizeof(POp) == 32, "two records per 64-byte scalar load");
constexpr int PL_NONE = 0, PL_REG = LOC_REG + 1, PL_THR = LOC_THR + 1, PL_OUT = LOC_OUT + 1;
constexpr int POP_NOP = 6;  // kind of the padding record: copies the amplitudes
// Factored 2x2 on register bit t0 (pool: r, s): t1 = 0 the unit-diagonal form
// [[1, r], [s, 1]], t1 = 1 the unit-antidiagonal form [[r, 1], [1, s]].  The
// diagonal factor the host split off stays in the phase's pending diagonal.
constexpr int POP_U1F = 7;
// A D1 record with location PL_NONE: the pending diagonal over the 4 register
// bits, one factor per register index (pool: 16 entries).

// Value of a non-register location bit for this thread / tile.
__device__ __forceinline__ uint32_t loc_val(int loc, int pos, uint32_t tid) {
  return loc == PL_THR ? ((tid >> pos) & 1u) : ((blockIdx.x >> pos) & 1u);
}

#define RQ_P2(F, ...)                      \
  switch (p0 * 4 + p1) {                   \
    case 1: F<0, 1>(__VA_ARGS__); break;   \
    case 2: F<0, 2>(__VA_ARGS__); break;   \
    case 3: F<0, 3>(__VA_ARGS__); break;   \
    case 4: F<1, 0>(__VA_ARGS__); break;   \
    case 6: F<1, 2>(__VA_ARGS__); break;   \
    case 7: F<1, 3>(__VA_ARGS__); break;   \
    case 8: F<2, 0>(__VA_ARGS__); break;   \
    case 9: F<2, 1>(__VA_ARGS__); break;   \
    case 11: F<2, 3>(__VA_ARGS__); break;  \
    case 12: F<3, 0>(__VA_ARGS__); break;  \
    case 13: F<3, 1>(__VA_ARGS__); break;  \
    default: F<3, 2>(__VA_ARGS__); break;  \
  }
#define RQ_P1(F, P, ...)                 \
  switch (P) {                           \
    case 0: F<0>(__VA_ARGS__); break;    \
    case 1: F<1>(__VA_ARGS__); break;    \
    case 2: F<2>(__VA_ARGS__); break;    \
    default: F<3>(__VA_ARGS__); break;   \
  }

// One op, x -> y.  U1 ops carry no controls (the planner folds only
// uncontrolled 1-qubit gates into them).  Factored 2x2 records, the most
// frequent kind, are tested first; every other field is decoded in the case
// that uses it.
template <class T>
__device__ __forceinline__ void apply_one(const Amps<T>& x, Amps<T>& y, const POp& r, const Cx<T>* __restrict__ pool,
                                          uint32_t tid) {
  const uint32_t w = r.x;
  const uint32_t kind = w & 7u;
  if (__builtin_expect(kind == (uint32_t)POP_U1F, 1)) {  // cloc/cpos: a tile-id bit that toggles the form
    const int t0 = (w >> 3) & 3u;
    // the toggle is a tile-id bit (never a thread bit): wave-uniform, a
    // scalar branch (integer arithmetic only, so the test stays on the
    // scalar unit)
    const uint32_t anti = __builtin_popcount(r.tmask & (blockIdx.x | 0x80000000u)) & 1u;
    T rr, ri, sr, si;
    if constexpr (sizeof(T) == 4) {
      rr = r.c[0]; ri = r.c[1]; sr = r.c[2]; si = r.c[3];
    } else {
      const Cx<T>* mp = pool + r.y;
      rr = mp[0].re; ri = mp[0].im; sr = mp[1].re; si = mp[1].im;
    }
    if (anti) {
      RQ_P1(u1f_anti, t0, x, y, rr, ri, sr, si)
    } else {
      RQ_P1(u1f, t0, x, y, rr, ri, sr, si)
    }
    return;
  }
  const Cx<T>* mp = pool + r.y;
  switch (kind) {
    case FOP_U1: {  // cloc/cpos of a 2x2: a tile-id bit that selects the second of two matrices
      const int t0 = (w >> 3) & 3u, cl = (w >> 7) & 3u, cp = (w >> 9) & 31u;
      RQ_P1(u1, t0, x, y, (cl == PL_OUT && loc_val(cl, cp, tid)) ? mp + 4 : mp)
      break;
    }
    case FOP_U2: {
      const int p0 = (w >> 3) & 3u, p1 = (w >> 5) & 3u;
      RQ_P2(u2, x, y, mp)
      break;
    }
    case FOP_CX: {
      const int t0 = (w >> 3) & 3u, cl = (w >> 7) & 3u, cp = (w >> 9) & 31u;
      if ((w >> 5) & 1u) {  // t1 = 1: relative phase on the target's set half first (register or lane control)
        Cx<T> rho;
        if constexpr (sizeof(T) == 4) rho = Cx<T>{r.c[0], r.c[1]};
        else rho = mp[0];
        if (cl == PL_REG) {
          const int p0 = t0, p1 = cp;
          RQ_P2(cx_rc_ph, x, y, rho)
        } else {
          RQ_P1(cx_sel_ph, t0, x, y, __builtin_amdgcn_ballot_w64(loc_val(cl, cp, tid) != 0u), rho)
        }
        break;
      }
      if (cl == PL_REG) {
        const int p0 = t0, p1 = cp;
        RQ_P2(cx_rc, x, y)
      } else if (cl == PL_NONE) {
        RQ_P1(cx_all, t0, x, y)
      } else {
        RQ_P1(cx_sel, t0, x, y, __builtin_amdgcn_ballot_w64(loc_val(cl, cp, tid) != 0u))
      }
      break;
    }
    case FOP_CZ: {
      const int dl0 = (w >> 14) & 3u, dp0 = (w >> 16) & 63u, dl1 = (w >> 22) & 3u, dp1 = (w >> 24) & 63u;
      if (dl0 == PL_REG && dl1 == PL_REG) {
        const int p0 = dp0, p1 = dp1;
        RQ_P2(cz_rr, x, y)
      } else if (dl0 == PL_REG) {
        RQ_P1(cz_r, dp0, x, y, loc_val(dl1, dp1, tid) << 31)
      } else if (dl1 == PL_REG) {
        RQ_P1(cz_r, dp1, x, y, loc_val(dl0, dp0, tid) << 31)
      } else {
        flip_all(x, y, (loc_val(dl0, dp0, tid) & loc_val(dl1, dp1, tid)) << 31);
      }
      break;
    }
    case FOP_D1: {
      const int dl0 = (w >> 14) & 3u, dp0 = (w >> 16) & 63u;
      if (dl0 == PL_REG) {
        RQ_P1(dsel_r, dp0, x, y, mp[0], mp[1])
      } else if (dl0 == PL_NONE) {
        dtab(x, y, mp);
      } else {
        dall(x, y, mp[loc_val(dl0, dp0, tid)]);
      }
      break;
    }
    case POP_NOP:
#pragma unroll
      for (int k = 0; k < FUSE_R; ++k) y[k] = fresh(x[k]);
      break;
    default: {  // FOP_D2
      const int dl0 = (w >> 14) & 3u, dp0 = (w >> 16) & 63u, dl1 = (w >> 22) & 3u, dp1 = (w >> 24) & 63u;
      if (dl0 == PL_REG && dl1 == PL_REG) {
        const int p0 = dp0, p1 = dp1;
        RQ_P2(d2_rr, x, y, mp)
      } else if (dl0 == PL_REG) {
        const uint32_t c1 = loc_val(dl1, dp1, tid);
        RQ_P1(dsel_r, dp0, x, y, mp[c1], mp[2 + c1])
      } else if (dl1 == PL_REG) {
        const uint32_t c0 = loc_val(dl0, dp0, tid);
        RQ_P1(dsel_r, dp1, x, y, mp[2 * c0], mp[2 * c0 + 1])
      } else {
        dall(x, y, mp[2 * loc_val(dl0, dp0, tid) + loc_val(dl1, dp1, tid)]);
      }
      break;
    }
  }
}

// Ops [b, e) of a phase, two per iteration (v -> w -> v).  Every phase holds
// an even number of records (pack_ops pads odd ones with POP_NOP): an odd
// tail would merge the two register sets after the loop, and the compiler
// then copies all 16 amplitudes on every iteration.
template <class T>
__device__ __forceinline__ void apply_ops(Amps<T>& v, const POp* __restrict__ ops, int b, int e,
                                          const Cx<T>* __restrict__ pool, uint32_t tid) {
  Cx<T> w[FUSE_R];
  for (int o = b; o < e; o += 2) {
    POp p0, p1;
    if constexpr (sizeof(T) == 4) {
      // both records in one 64-byte scalar load at the top of the iteration:
      // the inline coefficients of a factored 2x2 are then in registers when
      // its branch is taken, instead of a second load after the kind test
      typedef uint32_t u32x16 __attribute__((ext_vector_type(16)));
      const u32x16 raw = *reinterpret_cast<const u32x16*>(ops + o);
      p0.x = raw[0];
      p0.y = raw[1];
      p0.tmask = raw[6];
      p1.x = raw[8];
      p1.y = raw[9];
      p1.tmask = raw[14];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        p0.c[i] = __uint_as_float(raw[2 + i]);
        p1.c[i] = __uint_as_float(raw[10 + i]);
      }
    } else {
      p0 = ops[o];
      p1 = ops[o + 1];
    }
    apply_one(v, w, p0, pool, tid);
    apply_one(w, v, p1, pool, tid);
  }
}

// Matrix-core phase (complex f32).  The host composes all ops of the phase
// into one 16x16 complex unitary U over the register bits (one U per value of
// the phase's uniform condition bits) and stores it as the real 32x32 matrix
// [[Re U, -Im U], [Im U, Re U]] in the A-operand order of
// v_mfma_f32_32x32x2_f32.  The phase's mapping swaps the roles of register
// bit 3 and thread bit 5: lane bit 5 (h) holds register bit 3 and v[8c + j]
// is chain c's amplitude j | h << 3, so lanes l and l + 32 hold the two halves
// of one group of 16 amplitudes, 32 groups per chain and 2 chains per wave.
// Per chain, D = A X with X[K][n] = real K of group n: step s takes K = 2s + h,
// the real (s & 1 ? im : re) of amplitude (s >> 1) | h << 3, from lane n + 32h,
// and the output row order puts real (v & 1) of amplitude (v >> 1) | h << 3
// back in accumulator v of lane n + 32h.
//   mf_sel: 3 bytes, each kind:2 | pos:6 — MF_OUT: tile-id bit pos,
//   MF_WAVE: thread bit pos (6 or 7: the same in all lanes of a wave),
//   MF_CHAIN: thread bit 5 of the phase's VALU mapping (the chain index c).
// The variant index has bit i from byte i; variants are 1024 floats apart,
// float (s >> 2) * 256 + lane * 4 + (s & 3) of a variant is lane's A value at step s.
constexpr uint32_t MF_OUT = 1, MF_WAVE = 2, MF_CHAIN = 3;
constexpr int MF_FLOATS = 1024;

typedef float f32x16 __attribute__((ext_vector_type(16)));

// The A operand of a matrix-core phase's chain 0 (16 floats per lane, all
// four loads issued before the first product).  Chain 1 reuses it unless the
// unitary depends on the chain (cb != 0, wave-uniform); then chain 1's operand
// is loaded while chain 0 runs.  The chains run one after the other, so at
// most 32 operand registers are live (the two chains interleaved with both
// operands held 32 more, and the compiler then issued each load just before
// its products and waited for it).  Loading the operand before the phase's
// LDS gather made the kernel spill.
struct MfA {
  float a0[16];
  const float4* w1;
  uint32_t cb;
};

__device__ __forceinline__ void mf_load(float (&a)[16], const float4* __restrict__ w) {
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const float4 x = w[q * 64];
    a[4 * q] = x.x; a[4 * q + 1] = x.y; a[4 * q + 2] = x.z; a[4 * q + 3] = x.w;
  }
}

__device__ __forceinline__ void mf_fetch(MfA& m, const float* __restrict__ w, uint32_t sel, uint32_t tid) {
  uint32_t vb = 0, cb = 0;
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const uint32_t d = (sel >> (8 * i)) & 255u, kind = d >> 6, pos = d & 63u;
    if (kind == MF_OUT) vb |= ((blockIdx.x >> pos) & 1u) << i;
    else if (kind == MF_WAVE) vb |= ((tid >> pos) & 1u) << i;
    else if (kind == MF_CHAIN) cb = 1u << i;
  }
  vb = __builtin_amdgcn_readfirstlane(vb);
  const uint32_t lane = tid & 63u;
  mf_load(m.a0, reinterpret_cast<const float4*>(w + vb * MF_FLOATS) + lane);
  m.cb = cb;
  m.w1 = reinterpret_cast<const float4*>(w + (vb | cb) * MF_FLOATS) + lane;
}

// One chain: D = A X over the 8 amplitudes v[8c .. 8c + 7] of this lane.
template <int C>
__device__ __forceinline__ void mf_chain(Cx<float> (&v)[FUSE_R], const float (&a)[16]) {
  f32x16 d = {};
#pragma unroll
  for (int s = 0; s < 16; ++s) {
    const float x = (s & 1) ? v[8 * C + (s >> 1)].im : v[8 * C + (s >> 1)].re;
    d = __builtin_amdgcn_mfma_f32_32x32x2f32(a[s], x, d, 0, 0, 0);
  }
#pragma unroll
  for (int j = 0; j < 8; ++j) v[8 * C + j] = Cx<float>{d[2 * j], d[2 * j + 1]};
}

__device__ __forceinline__ void apply_mf(Cx<float> (&v)[FUSE_R], const MfA& m) {
  if (m.cb) {
    float a1[16];
    mf_load(a1, m.w1);
    mf_chain<0>(v, m.a0);
    mf_chain<1>(v, a1);
  } else {
    mf_chain<0>(v, m.a0);
    mf_chain<1>(v, m.a0);
  }
}

// Offsets of one phase mapping: thread part and per-register-bit parts of the
// tile index (treg) and of the global index (greg).
struct Map {
  uint32_t tthr;
  uint64_t gthr;
  uint32_t treg[FUSE_R_BITS];
  uint64_t greg[FUSE_R_BITS];
};

template <class T>
__device__ __forceinline__ Cx<T> lds_ld(const Cx<T>* lds, uint32_t byte) {
  return *reinterpret_cast<const Cx<T>*>(reinterpret_cast<const unsigned char*>(lds) + byte);
}
template <class T>
__device__ __forceinline__ void lds_st(Cx<T>* lds, uint32_t byte, const Cx<T>& x) {
  *reinterpret_cast<Cx<T>*>(reinterpret_cast<unsigned char*>(lds) + byte) = x;
}

// A phase record as three 64-byte scalar loads (dword d of the record is
// word(d)); every field the kernel reads is then in registers at once.
typedef int32_t i32x16 __attribute__((ext_vector_type(16)));
struct PhaseWords {
  i32x16 lo, hi, lr;  // lr: the register parts' LDS offsets (FPhase::lreg)
  __device__ __forceinline__ int32_t word(int d) const { return d < 16 ? lo[d] : hi[d - 16]; }
  __device__ __forceinline__ int32_t op_begin() const { return word(offsetof(FPhase, op_begin) / 4); }
  __device__ __forceinline__ int32_t op_end() const { return word(offsetof(FPhase, op_end) / 4); }
  __device__ __forceinline__ int32_t direct() const { return word(offsetof(FPhase, direct) / 4); }
  __device__ __forceinline__ int32_t mf_off() const { return word(offsetof(FPhase, mf_off) / 4); }
  __device__ __forceinline__ uint32_t mf_sel() const { return (uint32_t)word(offsetof(FPhase, mf_sel) / 4); }
  __device__ __forceinline__ int32_t regt(int b) const { return word(offsetof(FPhase, regt) / 4 + b); }
  __device__ __forceinline__ int32_t greg(int b) const { return word(offsetof(FPhase, greg) / 4 + b); }
  __device__ __forceinline__ int32_t thrt(int j) const { return word(offsetof(FPhase, thrt) / 4 + j); }
  __device__ __forceinline__ int32_t gthr(int j) const { return word(offsetof(FPhase, gthr) / 4 + j); }
  __device__ __forceinline__ uint32_t lreg(int k) const { return (uint32_t)lr[k]; }
};
static_assert(offsetof(FPhase, lreg) == 128, "the LDS offsets form the third 64-byte load");
// Complex f32 reads the LDS offsets too; the complex-f64 kernel computes
// them per access instead (16 more scalar registers would make it spill
// scalar registers throughout its op loop).
template <class T>
__device__ __forceinline__ PhaseWords load_phase(const FPhase* p) {
  const i32x16* v = reinterpret_cast<const i32x16*>(p);
  if constexpr (sizeof(T) == 4) return PhaseWords{v[0], v[1], v[2]};
  else return PhaseWords{v[0], v[1], i32x16{}};
}

// Branch-free: the block has 2^(m-4) threads, so thread-id bits past m-5 are
// zero, and the record's unused position entries are zero.
__device__ __forceinline__ Map make_map(const PhaseWords& ph, uint32_t tid, uint64_t outer) {
  Map mp;
  mp.tthr = 0;
  mp.gthr = outer;
#pragma unroll
  for (int j = 0; j < FUSE_MAX_M - FUSE_R_BITS; ++j) {
    const uint32_t b = (tid >> j) & 1u;
    mp.tthr |= b << (ph.thrt(j) & 31);
    mp.gthr |= (uint64_t)b << (ph.gthr(j) & 63);
  }
#pragma unroll
  for (int b = 0; b < FUSE_R_BITS; ++b) {
    mp.treg[b] = 1u << ph.regt(b);
    mp.greg[b] = 1ull << ph.greg(b);
  }
  return mp;
}

// The natural mapping: thread bits = tile bits 0..m-5, register bits = the
// top 4 tile bits.
__device__ __forceinline__ Map make_natural_map(const FStage& st, uint32_t tid, uint64_t outer) {
  Map mp;
  const int nt = st.m - FUSE_R_BITS;
  mp.tthr = tid;
  mp.gthr = outer;
  for (int j = 0; j < nt; ++j) mp.gthr |= (uint64_t)((tid >> j) & 1u) << st.S[j];
#pragma unroll
  for (int b = 0; b < FUSE_R_BITS; ++b) {
    mp.treg[b] = 1u << (nt + b);
    mp.greg[b] = 1ull << st.S[nt + b];
  }
  return mp;
}

template <int K>
__device__ __forceinline__ uint32_t treg_of(const Map& mp) {
  return ((K & 1) ? mp.treg[0] : 0u) | ((K & 2) ? mp.treg[1] : 0u) | ((K & 4) ? mp.treg[2] : 0u) |
         ((K & 8) ? mp.treg[3] : 0u);
}
template <int K>
__device__ __forceinline__ uint64_t greg_of(const Map& mp) {
  return ((K & 1) ? mp.greg[0] : 0ull) | ((K & 2) ? mp.greg[1] : 0ull) | ((K & 4) ? mp.greg[2] : 0ull) |
         ((K & 8) ? mp.greg[3] : 0ull);
}

// HBM accesses of a mapping: the thread's part and the (uniform) register
// part of an index have disjoint bits, so the address is the thread's
// pointer plus a uniform offset (one 64-bit add per access).
template <int K, class T>
struct Unroll {
  using V = typename Vec2<T>::type;
  static __device__ __forceinline__ void gather_g(Cx<T> (&v)[FUSE_R], const V* at, const Map& mp) {
    Unroll<K - 1, T>::gather_g(v, at, mp);
    const V x = at[greg_of<K - 1>(mp)];
    v[K - 1] = Cx<T>{x.x, x.y};
  }
  // |0..0>: 1 at global index 0, 0 elsewhere
  static __device__ __forceinline__ void zero_state(Cx<T> (&v)[FUSE_R], const Map& mp) {
    Unroll<K - 1, T>::zero_state(v, mp);
    v[K - 1] = Cx<T>{(mp.gthr | greg_of<K - 1>(mp)) == 0 ? T(1) : T(0), T(0)};
  }
  static __device__ __forceinline__ void scatter_g(const Cx<T> (&v)[FUSE_R], V* at, const Map& mp) {
    Unroll<K - 1, T>::scatter_g(v, at, mp);
    V x;
    x.x = v[K - 1].re;
    x.y = v[K - 1].im;
    at[greg_of<K - 1>(mp)] = x;
  }
  static __device__ __forceinline__ void gather_l(Cx<T> (&v)[FUSE_R], const Cx<T>* lds, const Map& mp,
                                                  uint32_t lt) {
    Unroll<K - 1, T>::gather_l(v, lds, mp, lt);
    v[K - 1] = lds_ld(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)));
  }
  // the same with the register parts' offsets from the phase record
  // (complex f32; see load_phase)
  static __device__ __forceinline__ void gather_lp(Cx<T> (&v)[FUSE_R], const Cx<T>* lds, const PhaseWords& ph,
                                                   const Map& mp, uint32_t lt) {
    Unroll<K - 1, T>::gather_lp(v, lds, ph, mp, lt);
    if constexpr (sizeof(T) == 4) v[K - 1] = lds_ld(lds, lt ^ ph.lreg(K - 1));
    else v[K - 1] = lds_ld(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)));
  }
  static __device__ __forceinline__ void scatter_lp(const Cx<T> (&v)[FUSE_R], Cx<T>* lds, const PhaseWords& ph,
                                                    const Map& mp, uint32_t lt) {
    Unroll<K - 1, T>::scatter_lp(v, lds, ph, mp, lt);
    if constexpr (sizeof(T) == 4) lds_st(lds, lt ^ ph.lreg(K - 1), v[K - 1]);
    else lds_st(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)), v[K - 1]);
  }
  static __device__ __forceinline__ void scatter_l(const Cx<T> (&v)[FUSE_R], Cx<T>* lds, const Map& mp,
                                                   uint32_t lt) {
    Unroll<K - 1, T>::scatter_l(v, lds, mp, lt);
    lds_st(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)), v[K - 1]);
  }
};
template <class T>
struct Unroll<0, T> {
  using V = typename Vec2<T>::type;
  static __device__ __forceinline__ void gather_g(Cx<T> (&)[FUSE_R], const V*, const Map&) {}
  static __device__ __forceinline__ void zero_state(Cx<T> (&)[FUSE_R], const Map&) {}
  static __device__ __forceinline__ void scatter_g(const Cx<T> (&)[FUSE_R], V*, const Map&) {}
  static __device__ __forceinline__ void gather_l(Cx<T> (&)[FUSE_R], const Cx<T>*, const Map&, uint32_t) {}
  static __device__ __forceinline__ void gather_lp(Cx<T> (&)[FUSE_R], const Cx<T>*, const PhaseWords&, const Map&,
                                                   uint32_t) {}
  static __device__ __forceinline__ void scatter_lp(const Cx<T> (&)[FUSE_R], Cx<T>*, const PhaseWords&, const Map&,
                                                    uint32_t) {}
  static __device__ __forceinline__ void scatter_l(const Cx<T> (&)[FUSE_R], Cx<T>*, const Map&, uint32_t) {}
};

template <class T>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(5))) void k_fused_stage(T* __restrict__ a, const FStage* __restrict__ stp,
                                                      const FPhase* __restrict__ phases, const POp* __restrict__ ops,
                                                      const Cx<T>* __restrict__ pool, int zero_in) {
  // A static array (the kernel's only LDS) sits at LDS address 0, so the
  // compiler folds its base into the addresses.
  __shared__ Cx<T> lds[1 << FUSE_MAX_M];
  using V2 = typename Vec2<T>::type;
  const FStage& st = *stp;
  const uint32_t tid = threadIdx.x;
  uint64_t outer = 0;
  {
    // tiles have at most 31 id bits: outer[0..31] in two 64-byte scalar
    // loads; bits of blockIdx.x past nouter are zero, so no test is needed
    const i32x16* ov = reinterpret_cast<const i32x16*>(st.outer);
    const i32x16 o0 = ov[0], o1 = ov[1];
    const uint64_t tile = blockIdx.x;
#pragma unroll
    for (int j = 0; j < 32; ++j) outer |= ((tile >> j) & 1ull) << ((j < 16 ? o0[j] : o1[j - 16]) & 63);
  }
  // All waves keep the default priority: raising it for the data movement
  // (HBM loads and stores, LDS transposes) once paid, and with the current
  // op records it costs 0.3-0.8 % (raising it for the ops instead, 3 %).
  Cx<T> v[FUSE_R];
  const int np = st.nphase;
  for (int p = 0; p < np; ++p) {
    const PhaseWords ph = load_phase<T>(phases + st.phase0 + p);
    const Map mp = make_map(ph, tid, outer);
    if (sizeof(T) == 4 && p == 0 && zero_in) {
      // the program's first stage on |0..0>: the amplitudes are made here
      // (no clearing pass before, no read of the state); complex f32 only
      // (in the complex-f64 kernel the branch costs scalar-register spills)
      const Map& m0 = mp;
      if (ph.direct()) {
        Unroll<FUSE_R, T>::zero_state(v, m0);
      } else {
        const Map nat = make_natural_map(st, tid, outer);
        Unroll<FUSE_R, T>::zero_state(v, nat);
        Unroll<FUSE_R, T>::scatter_l(v, lds, nat, lane_lds_addr<T>(nat.tthr));
        __syncthreads();
        Unroll<FUSE_R, T>::gather_lp(v, lds, ph, mp, lane_lds_addr<T>(mp.tthr));
        __syncthreads();
      }
    } else if (p == 0) {
      if (ph.direct()) {
        Unroll<FUSE_R, T>::gather_g(v, reinterpret_cast<const V2*>(a) + mp.gthr, mp);
      } else {
        const Map nat = make_natural_map(st, tid, outer);
        Unroll<FUSE_R, T>::gather_g(v, reinterpret_cast<const V2*>(a) + nat.gthr, nat);
        Unroll<FUSE_R, T>::scatter_l(v, lds, nat, lane_lds_addr<T>(nat.tthr));
        __syncthreads();
        Unroll<FUSE_R, T>::gather_lp(v, lds, ph, mp, lane_lds_addr<T>(mp.tthr));
        __syncthreads();
      }
    } else {
      Unroll<FUSE_R, T>::gather_lp(v, lds, ph, mp, lane_lds_addr<T>(mp.tthr));
      __syncthreads();
    }
    if constexpr (sizeof(T) == 4) {
      if (ph.mf_off() >= 0) {
        MfA mfa;
        mf_fetch(mfa, reinterpret_cast<const float*>(pool + ph.mf_off()), ph.mf_sel(), tid);
        apply_mf(v, mfa);
      } else {
        apply_ops(v, ops, ph.op_begin(), ph.op_end(), pool, tid);
      }
    } else {
      apply_ops(v, ops, ph.op_begin(), ph.op_end(), pool, tid);
    }
    if (p == np - 1) {
      if (ph.direct()) {
        Unroll<FUSE_R, T>::scatter_g(v, reinterpret_cast<V2*>(a) + mp.gthr, mp);
      } else {
        const Map nat = make_natural_map(st, tid, outer);
        Unroll<FUSE_R, T>::scatter_lp(v, lds, ph, mp, lane_lds_addr<T>(mp.tthr));
        __syncthreads();
        Unroll<FUSE_R, T>::gather_l(v, lds, nat, lane_lds_addr<T>(nat.tthr));
        Unroll<FUSE_R, T>::scatter_g(v, reinterpret_cast<V2*>(a) + nat.gthr, nat);
      }
    } else {
      Unroll<FUSE_R, T>::scatter_lp(v, lds, ph, mp, lane_lds_addr<T>(mp.tthr));
      __syncthreads();
    }
  }
}

// ---- matrix-core phases: host side ------------------------------------------

// VALU instructions per wave of one op (16 amplitudes per thread): a
// factored 2x2 is 8 pairs x 8 FMAs, a 4x4 8 x 16 x 4 / 2; diagonals one
// complex product per amplitude; permutations and sign flips one move per
// word.  A phase adds one 64-instruction pending-diagonal table.
int op_cost(const FOp& op) {
  switch (op.kind) {
    case FOP_U1: return 64;
    case FOP_U2: return 256;
    case FOP_D1:
    case FOP_D2: return 64;
    default: return 32;
  }
}

// The uniform condition bits of a matrix-core phase (at most 3).
struct MfBits {
  int n = 0;
  uint32_t kind[3] = {0, 0, 0}, pos[3] = {0, 0, 0};
  int find(uint32_t k, uint32_t p) const {
    for (int i = 0; i < n; ++i)
      if (kind[i] == k && pos[i] == p) return i;
    return -1;
  }
  bool add(uint32_t k, uint32_t p) {
    if (find(k, p) >= 0) return true;
    if (n == 3) return false;
    kind[n] = k;
    pos[n] = p;
    ++n;
    return true;
  }
  uint32_t sel() const {
    uint32_t s = 0;
    for (int i = 0; i < n; ++i) s |= ((kind[i] << 6) | pos[i]) << (8 * i);
    return s;
  }
};

// Uniform-bit key of a condition bit at (loc, pos) in a phase with nt thread
// bits: thread bit 5 -> the chain, 6 and 7 -> the wave, a global qubit ->
// its tile-id bit.  False when the bit differs between the 32 groups of a
// chain (thread bits 0..4) or is not located.
template <class TileBit>
bool mf_key(int loc, int pos, int nt, const TileBit& tile_bit, uint32_t* k, uint32_t* p) {
  if (loc == LOC_THR) {
    if (pos == 5) { *k = MF_CHAIN; *p = 5; return true; }
    if ((pos == 6 || pos == 7) && pos < nt) { *k = MF_WAVE; *p = (uint32_t)pos; return true; }
    return false;
  }
  if (loc == LOC_OUT) {
    const int j = tile_bit(pos);
    if (j < 0 || j >= 32) return false;
    *k = MF_OUT;
    *p = (uint32_t)j;
    return true;
  }
  return false;
}

// Collects the phase's uniform condition bits; false when the phase cannot
// run on the matrix cores.
template <class TileBit>
bool mf_phase_bits(const FusedPlan& plan, const FPhase& ph, int nt, const TileBit& tile_bit, MfBits* bits) {
  if (nt < 6) return false;
  for (int o = ph.op_begin; o < ph.op_end; ++o) {
    const FOp& op = plan.ops[o];
    uint32_t k, p;
    if (op.kind == FOP_U1 || op.kind == FOP_U2) {
      if (op.ctl_reg || op.ctl_thr || op.ctl_out) return false;
    } else if (op.kind == FOP_CX) {
      const int nctl = __builtin_popcount(op.ctl_reg) + __builtin_popcount(op.ctl_thr) + __builtin_popcountll(op.ctl_out);
      if (nctl > 1) return false;
      if (op.ctl_thr && !(mf_key(LOC_THR, __builtin_ctz(op.ctl_thr), nt, tile_bit, &k, &p) && bits->add(k, p)))
        return false;
      if (op.ctl_out && !(mf_key(LOC_OUT, __builtin_ctzll(op.ctl_out), nt, tile_bit, &k, &p) && bits->add(k, p)))
        return false;
    } else {
      if (op.dloc0 != LOC_REG && !(mf_key(op.dloc0, op.dpos0, nt, tile_bit, &k, &p) && bits->add(k, p))) return false;
      if (op.kind != FOP_D1 && op.dloc1 != LOC_REG &&
          !(mf_key(op.dloc1, op.dpos1, nt, tile_bit, &k, &p) && bits->add(k, p)))
        return false;
    }
  }
  return true;
}

// U[r][c] (row = output register index) of the phase's ops, in order, with
// variant vi's values of the uniform bits.
template <class TileBit>
void mf_compose(const FusedPlan& plan, const FPhase& ph, int nt, const TileBit& tile_bit, const MfBits& bits,
                uint32_t vi, rq_c64 (*U)[FUSE_R]) {
  auto bitv = [&](int loc, int pos, int k) -> int {
    if (loc == LOC_REG) return (k >> pos) & 1;
    uint32_t kk = 0, pp = 0;
    mf_key(loc, pos, nt, tile_bit, &kk, &pp);
    return (int)((vi >> bits.find(kk, pp)) & 1u);
  };
  auto mul = [](rq_c64 a, rq_c64 b) { return rq_c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; };
  auto add = [](rq_c64 a, rq_c64 b) { return rq_c64{a.re + b.re, a.im + b.im}; };
  for (int col = 0; col < FUSE_R; ++col) {
    rq_c64 x[FUSE_R];
    for (int k = 0; k < FUSE_R; ++k) x[k] = rq_c64{k == col ? 1.0 : 0.0, 0.0};
    for (int o = ph.op_begin; o < ph.op_end; ++o) {
      const FOp& op = plan.ops[o];
      const rq_c64* mp = &plan.pool[op.mat];
      if (op.kind == FOP_U1) {
        const int bb = 1 << op.t0;
        for (int k = 0; k < FUSE_R; ++k) {
          if (k & bb) continue;
          const rq_c64 a = x[k], b = x[k | bb];
          x[k] = add(mul(mp[0], a), mul(mp[1], b));
          x[k | bb] = add(mul(mp[2], a), mul(mp[3], b));
        }
      } else if (op.kind == FOP_U2) {
        const int b0 = 1 << op.t0, b1 = 1 << op.t1;
        for (int k = 0; k < FUSE_R; ++k) {
          if (k & (b0 | b1)) continue;
          const int idx[4] = {k, k | b1, k | b0, k | b0 | b1};
          const rq_c64 in[4] = {x[idx[0]], x[idx[1]], x[idx[2]], x[idx[3]]};
          for (int r = 0; r < 4; ++r) {
            rq_c64 acc{0.0, 0.0};
            for (int c = 0; c < 4; ++c) acc = add(acc, mul(mp[r * 4 + c], in[c]));
            x[idx[r]] = acc;
          }
        }
      } else if (op.kind == FOP_CX) {
        if (op.ctl_thr && !bitv(LOC_THR, __builtin_ctz(op.ctl_thr), 0)) continue;
        if (op.ctl_out && !bitv(LOC_OUT, __builtin_ctzll(op.ctl_out), 0)) continue;
        const int bb = 1 << op.t0;
        for (int k = 0; k < FUSE_R; ++k) {
          if ((k & bb) || (k & (int)op.ctl_reg) != (int)op.ctl_reg) continue;
          std::swap(x[k], x[k | bb]);
        }
      } else if (op.kind == FOP_CZ) {
        for (int k = 0; k < FUSE_R; ++k)
          if (bitv(op.dloc0, op.dpos0, k) & bitv(op.dloc1, op.dpos1, k)) x[k] = rq_c64{-x[k].re, -x[k].im};
      } else if (op.kind == FOP_D1) {
        for (int k = 0; k < FUSE_R; ++k) x[k] = mul(mp[bitv(op.dloc0, op.dpos0, k)], x[k]);
      } else {  // FOP_D2
        for (int k = 0; k < FUSE_R; ++k)
          x[k] = mul(mp[2 * bitv(op.dloc0, op.dpos0, k) + bitv(op.dloc1, op.dpos1, k)], x[k]);
      }
    }
    for (int r = 0; r < FUSE_R; ++r) U[r][col] = x[r];
  }
}

// A value of lane `lane` at MFMA step s: row m = lane % 32 and column
// K = 2s + lane / 32 of the real matrix of U, in the row / column order of
// apply_mf.
void mf_operands(const rq_c64 (*U)[FUSE_R], float* w) {
  for (int m = 0; m < 32; ++m) {
    const int h = (m >> 2) & 1, vv = ((m >> 3) << 2) | (m & 3);
    const int ao = (vv >> 1) | (h << 3), po = vv & 1;
    for (int K = 0; K < 32; ++K) {
      const int hk = K & 1, s = K >> 1;
      const int ai = (s >> 1) | (hk << 3), pi = s & 1;
      const rq_c64 u = U[ao][ai];
      const double val = po == 0 ? (pi == 0 ? u.re : -u.im) : (pi == 0 ? u.im : u.re);
      const int lane = m + 32 * hk;
      w[(s >> 2) * 256 + lane * 4 + (s & 3)] = (float)val;
    }
  }
}

// Pending diagonal of a VALU phase over its 4 register bits: entry k is
// g[k] * d[0][k_0] * d[1][k_1] * d[2][k_2] * d[3][k_3].  The per-bit factors d
// come from the factored 2x2 ops and from 1-qubit diagonals; g holds what
// couples bits (CZ signs, 2-qubit diagonals, factors carried through a
// CNOT's permutation).  g is updated only by multiplications whose factor is
// a function of index bits and by permutations, so two entries that differ in
// bits g does not depend on hold bit-identical values: `==` tests
// independence exactly.
struct PendingDiag {
  rq_c64 d[FUSE_R_BITS][2];
  rq_c64 g[FUSE_R];
  PendingDiag() {
    for (auto& b : d) b[0] = b[1] = rq_c64{1.0, 0.0};
    for (auto& x : g) x = rq_c64{1.0, 0.0};
  }
  static rq_c64 mul(rq_c64 a, rq_c64 b) { return rq_c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
  static bool eq(rq_c64 a, rq_c64 b) { return a.re == b.re && a.im == b.im; }
  bool identity() const {
    const rq_c64 one{1.0, 0.0};
    for (const auto& b : d)
      if (!eq(b[0], one) || !eq(b[1], one)) return false;
    for (const auto& x : g)
      if (!eq(x, one)) return false;
    return true;
  }
  // g constant over the 4 bits (its entries all equal): the constant moves
  // into d[0] and g becomes 1; false when g couples bits.
  bool separate() {
    for (int k = 1; k < FUSE_R; ++k)
      if (!eq(g[k], g[0])) return false;
    d[0][0] = mul(d[0][0], g[0]);
    d[0][1] = mul(d[0][1], g[0]);
    for (auto& x : g) x = rq_c64{1.0, 0.0};
    return true;
  }
  rq_c64 entry(int k) const {
    rq_c64 v = g[k];
    for (int b = 0; b < FUSE_R_BITS; ++b) v = mul(v, d[b][(k >> b) & 1]);
    return v;
  }
  bool free_of(int b) const {  // g does not depend on bit b
    for (int k = 0; k < FUSE_R; ++k)
      if (!eq(g[k], g[k ^ (1 << b)])) return false;
    return true;
  }
  bool flip_invariant(int b) const { return free_of(b) && eq(d[b][0], d[b][1]); }
  void clear_bit(int b) { d[b][0] = d[b][1] = rq_c64{1.0, 0.0}; }
  void set_bit(int b, const rq_c64 e[2]) { d[b][0] = e[0]; d[b][1] = e[1]; }
  // M diag(d[b]): the factor applied before M moves into its columns.
  void fold_columns(int b, rq_c64 m[4]) const {
    m[0] = mul(m[0], d[b][0]); m[2] = mul(m[2], d[b][0]);
    m[1] = mul(m[1], d[b][1]); m[3] = mul(m[3], d[b][1]);
  }
  // 4x4 on (b0 = row-high bit, b1): column c = 2 bit_b0 + bit_b1.
  void fold_columns4(int b0, int b1, rq_c64 m[16]) const {
    for (int r = 0; r < 4; ++r)
      for (int c = 0; c < 4; ++c) m[r * 4 + c] = mul(mul(m[r * 4 + c], d[b0][c >> 1]), d[b1][c & 1]);
  }
  // CNOT(c -> t) after the diagonal: D then X_t^c equals X_t^c then D o pi.
  void cx_reg(int c, int t) {
    for (int k = 0; k < FUSE_R; ++k) g[k] = mul(g[k], d[t][(k >> t) & 1]);
    clear_bit(t);
    rq_c64 h[FUSE_R];
    for (int k = 0; k < FUSE_R; ++k) h[k] = g[k ^ (((k >> c) & 1) << t)];
    for (int k = 0; k < FUSE_R; ++k) g[k] = h[k];
  }
  void cx_all(int t) {
    std::swap(d[t][0], d[t][1]);
    rq_c64 h[FUSE_R];
    for (int k = 0; k < FUSE_R; ++k) h[k] = g[k ^ (1 << t)];
    for (int k = 0; k < FUSE_R; ++k) g[k] = h[k];
  }
  void cz(int p0, int p1) {
    for (int k = 0; k < FUSE_R; ++k)
      if ((k >> p0) & (k >> p1) & 1) g[k] = rq_c64{-g[k].re, -g[k].im};
  }
  void d1(int b, const rq_c64* f) { d[b][0] = mul(d[b][0], f[0]); d[b][1] = mul(d[b][1], f[1]); }
  void d2(int p0, int p1, const rq_c64* f) {
    for (int k = 0; k < FUSE_R; ++k) g[k] = mul(g[k], f[2 * ((k >> p0) & 1) + ((k >> p1) & 1)]);
  }
};

// M = diag(e0, e1) F with F = [[1, r], [s, 1]] (returns 0) or [[r, 1], [1, s]]
// (returns 1), the form whose ratios are the smaller (|r|, |s| <= 1 for a
// unitary M); -1 when neither exists.
int factor_2x2(const rq_c64 m[4], rq_c64 rs[2], rq_c64 e[2]) {
  auto mag2 = [](rq_c64 a) { return a.re * a.re + a.im * a.im; };
  auto div = [&](rq_c64 a, rq_c64 b) {
    const double n = mag2(b);
    return rq_c64{(a.re * b.re + a.im * b.im) / n, (a.im * b.re - a.re * b.im) / n};
  };
  const double dg = mag2(m[0]) * mag2(m[3]), ad = mag2(m[1]) * mag2(m[2]);
  if (dg >= ad && dg > 0.0) {  // diag(m00, m11) [[1, m01/m00], [m10/m11, 1]]
    rs[0] = div(m[1], m[0]);
    rs[1] = div(m[2], m[3]);
    e[0] = m[0];
    e[1] = m[3];
    return 0;
  }
  if (ad > 0.0) {  // diag(m01, m10) [[m00/m01, 1], [1, m11/m10]]
    rs[0] = div(m[0], m[1]);
    rs[1] = div(m[3], m[2]);
    e[0] = m[1];
    e[1] = m[2];
    return 1;
  }
  return -1;
}

// True when op `x` involves register bit b in any role.
bool touches_reg_bit(const FOp& x, int b) {
  switch (x.kind) {
    case FOP_U1: return x.t0 == b;
    case FOP_U2: return x.t0 == b || x.t1 == b;
    case FOP_CX: return x.t0 == b || ((x.ctl_reg >> b) & 1u);
    default:  // D1, D2, CZ
      return (x.dloc0 == LOC_REG && x.dpos0 == b) ||
             (x.kind != FOP_D1 && x.dloc1 == LOC_REG && x.dpos1 == b);
  }
}

// Packs the plan's ops for the kernel: one 8-byte record each, control and
// diagonal positions of outer qubits turned into tile-id (blockIdx) bits, and
// every phase padded to an even record count with POP_NOP.  `phases` receives
// the plan's phases with op ranges into `out`; `pool` the matrix pool,
// extended by the folded matrices below.
//
// Folding: a CNOT whose target is register bit b and whose control is absent
// or a tile-id bit c (X_b, or X_b when c = 1) commutes with every op that does
// not involve bit b, and c is constant over the tile.  Such an op is moved to
// the next op of its phase that involves b when that op is an uncontrolled 2x2
// on b (M), and merged into it: M X is exact (a column swap), and with a
// control bit the 2x2 takes the pair [M, M X] and picks by bit c.
//
// Factoring: every 2x2 M that reaches the emitter is split as diag(e0, e1) F
// with F = [[1, r], [s, 1]] or [[r, 1], [1, s]] (factor_2x2), and only F is
// applied (8 FMAs per amplitude pair instead of 16).  The split-off diagonal
// joins the phase's pending diagonal (PendingDiag), which also absorbs CZ and
// diagonal ops on register bits, moves through CNOTs as a permutation, folds
// into the columns of the next 2x2 / 4x4 on its bits, and is applied as one
// 16-entry table record where it cannot move further (a CNOT with a per-lane
// condition on a bit it depends on, an op on a bit it couples to others) and
// at the end of the phase.
//
// Carried factors: a phase whose pending diagonal holds only per-bit factors
// (g constant) applies no table at its end.  The factor of register bit b
// (global qubit q) is carried to the next phase in which q is a register bit,
// where it starts that phase's pending factor of q (a matrix-core phase folds
// it into the columns of its unitaries).  In between, every op that involves
// q acts on it as a control or a diagonal bit, and those commute with a
// diagonal on q.  A qubit's last register phase of the program applies its
// factor there.
//
// Matrix-core phases (mf_min >= 0, complex f32): a phase whose ops cost at
// least mf_min VALU instructions per wave and whose conditions are register
// bits or uniform bits (MfBits) gets no records: its composed unitaries go to
// the pool (aligned to 256 bytes) and its mapping swaps register bit 3 with
// thread bit 5 (the layout apply_mf expects).
int pack_ops(const FusedPlan& plan, int mf_min, std::vector<POp>& out, std::vector<FPhase>& phases,
             std::vector<rq_c64>& pool) {
  out.clear();
  out.reserve(plan.ops.size() + plan.phases.size());
  phases = plan.phases;
  pool = plan.pool;
  auto tile_bit_of = [](const FStage& st) {
    return [&st](int q) {
      for (int j = 0; j < st.nouter; ++j)
        if (st.outer[j] == q) return j;
      return -1;
    };
  };
  std::vector<char> is_mf(plan.phases.size(), 0);
  std::vector<MfBits> mfb(plan.phases.size());
  if (mf_min >= 0) {
    for (const FStage& st : plan.stages) {
      const auto tile_bit = tile_bit_of(st);
      const int nt = st.m - FUSE_R_BITS;
      for (int ph = st.phase0; ph < st.phase0 + st.nphase; ++ph) {
        const FPhase& f = plan.phases[ph];
        if (f.op_end == f.op_begin) continue;
        int cost = 64 + (((f.op_end - f.op_begin) & 1) ? 32 : 0);
        for (int o = f.op_begin; o < f.op_end; ++o) cost += op_cost(plan.ops[o]);
        if (cost < mf_min) continue;
        MfBits b;
        if (!mf_phase_bits(plan, f, nt, tile_bit, &b)) continue;
        is_mf[ph] = 1;
        mfb[ph] = b;
      }
    }
  }
  // last phase in which each global qubit is a register bit; the factors carried between phases
  std::vector<int> last_reg(64, -1);  // global qubits are < 64
  for (const FStage& st : plan.stages)
    for (int ph = st.phase0; ph < st.phase0 + st.nphase; ++ph)
      for (int b = 0; b < FUSE_R_BITS; ++b) last_reg[st.S[plan.phases[ph].regt[b]]] = ph;
  std::vector<std::array<rq_c64, 2>> carry(64, {rq_c64{1.0, 0.0}, rq_c64{1.0, 0.0}});
  auto qubit_of = [&](const FStage& st, int ph, int b) { return st.S[plan.phases[ph].regt[b]]; };
  std::vector<int> fold_into(plan.ops.size(), -1);  // op -> index of the 2x2 it merges into
  std::vector<int> folded_by(plan.ops.size(), -1);  // 2x2 -> the op merged into it
  for (size_t pi = 0; pi < plan.phases.size(); ++pi) {
    if (is_mf[pi]) continue;
    const FPhase& ph = plan.phases[pi];
    for (int o = ph.op_begin; o < ph.op_end; ++o) {
      const FOp& op = plan.ops[o];
      int b = -1;
      if (op.kind == FOP_CX && !op.ctl_reg && !op.ctl_thr) b = op.t0;
      if (b < 0) continue;
      for (int q = o + 1; q < ph.op_end; ++q) {
        const FOp& x = plan.ops[q];
        if (!touches_reg_bit(x, b)) continue;
        if (x.kind == FOP_U1 && x.t0 == b && !x.ctl_reg && !x.ctl_thr && !x.ctl_out && folded_by[q] < 0 &&
            fold_into[q] < 0) {
          fold_into[o] = q;
          folded_by[q] = o;
        }
        break;
      }
    }
  }
  for (const FStage& st : plan.stages) {
    const auto tile_bit = tile_bit_of(st);
    auto enc = [&](int loc, int pos, uint32_t* l, uint32_t* p) {
      if (loc == LOC_OUT) {
        const int j = tile_bit(pos);
        if (j < 0) return false;
        pos = j;
      }
      *l = (uint32_t)loc + 1u;
      *p = (uint32_t)pos;
      return pos >= 0 && pos < 32;
    };
    for (int ph = st.phase0; ph < st.phase0 + st.nphase; ++ph) {
      phases[ph].op_begin = (int32_t)out.size();
      if (is_mf[ph]) {
        while (pool.size() % 32) pool.push_back(rq_c64{0.0, 0.0});
        const size_t off = pool.size(), nv = (size_t)1 << mfb[ph].n;
        if (off + nv * (MF_FLOATS / 2) >= ((size_t)1 << 31)) return fail(RQ_ERR_ARGUMENT, "fused matrix pool too large");
        pool.resize(off + nv * (MF_FLOATS / 2), rq_c64{0.0, 0.0});
        std::vector<float> w(MF_FLOATS);
        for (uint32_t vi = 0; vi < nv; ++vi) {
          rq_c64 U[FUSE_R][FUSE_R];
          mf_compose(plan, plan.phases[ph], st.m - FUSE_R_BITS, tile_bit, mfb[ph], vi, U);
          for (int col = 0; col < FUSE_R; ++col) {  // the carried factors act first: U diag(carry)
            rq_c64 f{1.0, 0.0};
            for (int b = 0; b < FUSE_R_BITS; ++b) f = PendingDiag::mul(f, carry[qubit_of(st, ph, b)][(col >> b) & 1]);
            for (int r = 0; r < FUSE_R; ++r) U[r][col] = PendingDiag::mul(U[r][col], f);
          }
          mf_operands(U, w.data());
          for (int f = 0; f < MF_FLOATS; ++f) {
            rq_c64& e = pool[off + vi * (MF_FLOATS / 2) + f / 2];
            (f & 1 ? e.im : e.re) = (double)w[f];
          }
        }
        for (int b = 0; b < FUSE_R_BITS; ++b) carry[qubit_of(st, ph, b)] = {rq_c64{1.0, 0.0}, rq_c64{1.0, 0.0}};
        phases[ph].op_end = (int32_t)out.size();
        phases[ph].mf_off = (int32_t)off;
        phases[ph].mf_sel = mfb[ph].sel();
        std::swap(phases[ph].regt[3], phases[ph].thrt[5]);
        continue;
      }
      PendingDiag pd;
      for (int b = 0; b < FUSE_R_BITS; ++b) {  // factors carried in from earlier phases
        auto& c = carry[qubit_of(st, ph, b)];
        pd.set_bit(b, c.data());
        c = {rq_c64{1.0, 0.0}, rq_c64{1.0, 0.0}};
      }
      auto emit = [&](uint32_t kind, uint32_t t0, uint32_t t1, uint32_t cl, uint32_t cp, uint32_t dl0, uint32_t dp0,
                      uint32_t dl1, uint32_t dp1, uint32_t mat) {
        POp w{};
        w.x = kind | (t0 << 3) | (t1 << 5) | (cl << 7) | (cp << 9) | (dl0 << 14) | (dp0 << 16) | (dl1 << 22) |
              (dp1 << 24);
        w.y = mat;
        if (kind == (uint32_t)POP_U1F) w.tmask = (t1 << 31) | (cl == (uint32_t)PL_OUT ? 1u << cp : 0u);
        out.push_back(w);
      };
      auto flush = [&]() {  // apply the pending diagonal as one table record
        if (pd.identity()) return;
        const uint32_t mat = (uint32_t)pool.size();
        for (int k = 0; k < FUSE_R; ++k) pool.push_back(pd.entry(k));
        emit((uint32_t)FOP_D1, 0, 0, 0, 0, (uint32_t)PL_NONE, 0, 0, 0, mat);
        pd = PendingDiag();
      };
      // True when an op after o in this phase is a 2x2 or 4x4 on register bit b0 or b1.
      auto dense_later = [&](int o, int b0, int b1) {
        for (int q = o + 1; q < plan.phases[ph].op_end; ++q) {
          const FOp& x = plan.ops[q];
          if (fold_into[q] >= 0) continue;
          if (x.kind == FOP_U1 && (x.t0 == b0 || x.t0 == b1)) return true;
          if (x.kind == FOP_U2 && (x.t0 == b0 || x.t0 == b1 || x.t1 == b0 || x.t1 == b1)) return true;
        }
        return false;
      };
      for (int o = plan.phases[ph].op_begin; o < plan.phases[ph].op_end; ++o) {
        const FOp& op = plan.ops[o];
        if (fold_into[o] >= 0) continue;  // merged into a later 2x2
        uint32_t cl = 0, cp = 0, dl0 = 0, dp0 = 0, dl1 = 0, dp1 = 0;
        uint32_t mat = (uint32_t)op.mat;
        if (folded_by[o] >= 0) {
          const FOp& f = plan.ops[folded_by[o]];  // a CNOT
          const rq_c64* m = &plan.pool[op.mat];
          const rq_c64 mf[4] = {m[1], m[0], m[3], m[2]};  // M X: columns swapped
          const int cq = f.ctl_out ? __builtin_ctzll(f.ctl_out) : -1;  // the condition's global qubit (none: always)
          mat = (uint32_t)pool.size();
          if (cq < 0) {
            pool.insert(pool.end(), mf, mf + 4);
          } else {
            const int j = tile_bit(cq);
            if (j < 0 || j >= 32) return fail(RQ_ERR_ARGUMENT, "fused control outside the tile id");
            const rq_c64 m0[4] = {m[0], m[1], m[2], m[3]};
            pool.insert(pool.end(), m0, m0 + 4);
            pool.insert(pool.end(), mf, mf + 4);
            cl = PL_OUT;
            cp = (uint32_t)j;
          }
          if (pool.size() >= (size_t(1) << 32)) return fail(RQ_ERR_ARGUMENT, "fused matrix pool too large");
        }
        if (op.kind == FOP_U1 && (op.ctl_reg || op.ctl_thr || op.ctl_out))
          return fail(RQ_ERR_ARGUMENT, "fused 1-qubit op with a control");
        if (op.kind == FOP_CX) {
          const int nctl = __builtin_popcount(op.ctl_reg) + __builtin_popcount(op.ctl_thr) +
                           __builtin_popcountll(op.ctl_out);
          if (nctl > 1) return fail(RQ_ERR_ARGUMENT, "fused op with more than one control");
          if (op.ctl_reg) { cl = PL_REG; cp = (uint32_t)__builtin_ctz(op.ctl_reg); }
          else if (op.ctl_thr) { cl = PL_THR; cp = (uint32_t)__builtin_ctz(op.ctl_thr); }
          else if (op.ctl_out) {
            const int j = tile_bit(__builtin_ctzll(op.ctl_out));
            if (j < 0) return fail(RQ_ERR_ARGUMENT, "fused control outside the tile id");
            cl = PL_OUT;
            cp = (uint32_t)j;
          }
        }
        if (op.kind == FOP_D1 || op.kind == FOP_D2 || op.kind == FOP_CZ) {
          if (!enc(op.dloc0, op.dpos0, &dl0, &dp0)) return fail(RQ_ERR_ARGUMENT, "fused diagonal bit not located");
          if (op.kind != FOP_D1 && !enc(op.dloc1, op.dpos1, &dl1, &dp1))
            return fail(RQ_ERR_ARGUMENT, "fused diagonal bit not located");
        }
        const int b = op.t0;
        switch (op.kind) {
          case FOP_U1: {
            if (!pd.free_of(b)) flush();
            // With cl = PL_OUT the pool holds the pair [M, M X] selected by a tile-id bit (a CNOT
            // controlled by it was folded in), which acts after the pending factor on b: the pair is
            // [M diag(d), M X diag(d)].  When d[0] = d[1] that is [M', M' X] with M' = M diag(d) =
            // diag(e) F, and F X is F's other unit form with the same r and s: the tile-id bit toggles
            // the form.  Otherwise each matrix takes the factor and the pair runs unfactored.
            if (cl == PL_OUT && !pd.flip_invariant(b)) {
              const uint32_t nm = (uint32_t)pool.size();
              for (int v = 0; v < 2; ++v) {
                rq_c64 m2[4] = {pool[mat + 4 * v], pool[mat + 4 * v + 1], pool[mat + 4 * v + 2], pool[mat + 4 * v + 3]};
                pd.fold_columns(b, m2);
                pool.insert(pool.end(), m2, m2 + 4);
              }
              pd.clear_bit(b);
              emit((uint32_t)FOP_U1, (uint32_t)b, 0, cl, cp, 0, 0, 0, 0, nm);
              break;
            }
            rq_c64 m2[4] = {pool[mat], pool[mat + 1], pool[mat + 2], pool[mat + 3]};
            pd.fold_columns(b, m2);
            pd.clear_bit(b);
            rq_c64 rs[2], e[2];
            int form = factor_2x2(m2, rs, e);
            const uint32_t nm = (uint32_t)pool.size();
            if (form < 0) {  // a zero row or column: no unit form, apply the whole matrix (pair)
              rq_c64 mx[4] = {m2[1], m2[0], m2[3], m2[2]};
              pool.insert(pool.end(), m2, m2 + 4);
              pool.insert(pool.end(), mx, mx + 4);
              emit((uint32_t)FOP_U1, (uint32_t)b, 0, cl, cp, 0, 0, 0, 0, nm);
            } else {
              pool.insert(pool.end(), rs, rs + 2);
              emit((uint32_t)POP_U1F, (uint32_t)b, (uint32_t)form, cl, cp, 0, 0, 0, 0, nm);
              pd.set_bit(b, e);
            }
            break;
          }
          case FOP_U2: {
            if (!pd.free_of(op.t0) || !pd.free_of(op.t1)) flush();
            rq_c64 m4[16];
            for (int k = 0; k < 16; ++k) m4[k] = pool[mat + k];
            pd.fold_columns4(op.t0, op.t1, m4);
            pd.clear_bit(op.t0);
            pd.clear_bit(op.t1);
            const uint32_t nm = (uint32_t)pool.size();
            pool.insert(pool.end(), m4, m4 + 16);
            emit((uint32_t)FOP_U2, (uint32_t)op.t0, (uint32_t)op.t1, 0, 0, 0, 0, 0, 0, nm);
            break;
          }
          case FOP_CX: {
            // A controlled CNOT after a pending factor (d0, d1) on its target: the relative phase
            // d1 / d0 goes into the record (applied to the target's set half before the
            // permutation) and d0, the same for both values of the bit, stays pending.  Moved
            // through the CNOT, (d0, d1) would couple the target to the control (a table flush).
            uint32_t phase = 0;
            if (cl != PL_NONE && pd.free_of(b) && !PendingDiag::eq(pd.d[b][0], pd.d[b][1]) &&
                (pd.d[b][0].re != 0.0 || pd.d[b][0].im != 0.0)) {
              const rq_c64 d0 = pd.d[b][0], d1 = pd.d[b][1];
              const double n = d0.re * d0.re + d0.im * d0.im;
              mat = (uint32_t)pool.size();
              pool.push_back(rq_c64{(d1.re * d0.re + d1.im * d0.im) / n, (d1.im * d0.re - d1.re * d0.im) / n});
              pd.d[b][1] = d0;
              phase = 1;
            }
            if (cl == PL_REG) pd.cx_reg((int)cp, b);
            else if (cl == PL_NONE) pd.cx_all(b);
            else if (!pd.flip_invariant(b)) flush();  // a per-lane / per-tile condition
            emit((uint32_t)FOP_CX, (uint32_t)b, phase, cl, cp, 0, 0, 0, 0, mat);
            break;
          }
          case FOP_CZ:
            // Absorbed into the pending diagonal unless a later 2x2 / 4x4 of the phase acts on one of its
            // bits: the coupling would force a table flush (64 instructions) where the CZ record costs 32.
            if (dl0 == PL_REG && dl1 == PL_REG && !dense_later(o, (int)dp0, (int)dp1)) pd.cz((int)dp0, (int)dp1);
            else emit((uint32_t)FOP_CZ, 0, 0, 0, 0, dl0, dp0, dl1, dp1, mat);
            break;
          case FOP_D1:
            if (dl0 == PL_REG) pd.d1((int)dp0, &pool[mat]);  // absorbed
            else emit((uint32_t)FOP_D1, 0, 0, 0, 0, dl0, dp0, 0, 0, mat);
            break;
          default:  // FOP_D2
            if (dl0 == PL_REG && dl1 == PL_REG) pd.d2((int)dp0, (int)dp1, &pool[mat]);  // absorbed
            else emit((uint32_t)FOP_D2, 0, 0, 0, 0, dl0, dp0, dl1, dp1, mat);
            break;
        }
      }
      if (pd.separate()) {  // per-bit factors only: carry those of qubits with a later register phase
        for (int b = 0; b < FUSE_R_BITS; ++b) {
          const int q = qubit_of(st, ph, b);
          if (last_reg[q] == ph) continue;
          carry[q] = {pd.d[b][0], pd.d[b][1]};
          pd.clear_bit(b);
        }
      }
      flush();
      if ((out.size() - (size_t)phases[ph].op_begin) & 1) out.push_back(POp{(uint32_t)POP_NOP, 0u, {0.f, 0.f, 0.f, 0.f}, 0u, 0u});
      phases[ph].op_end = (int32_t)out.size();
    }
  }
  for (const FStage& st : plan.stages)  // global qubits of each phase's register and thread bits
    for (int ph = st.phase0; ph < st.phase0 + st.nphase; ++ph) {
      FPhase& f = phases[ph];
      for (int b = 0; b < FUSE_R_BITS; ++b) f.greg[b] = st.S[f.regt[b]];
      for (int j = 0; j < st.m - FUSE_R_BITS; ++j) f.gthr[j] = st.S[f.thrt[j]];
    }
  return RQ_OK;
}

}  // namespace

// The packed program of the last fused segment a state ran, kept in s.prog:
// the exact gates and knobs it was planned from, and where its parts sit.
struct FusedProgram {
  bool valid = false;
  std::vector<rq_gate> gates;
  uint32_t nq = 0;
  int prec = -1;
  FusionConfig cfg;
  std::vector<FStage> stages;
  size_t o_st = 0, o_ph = 0, o_op = 0, o_pool = 0;
  uint64_t unitary = 0;
};

namespace {

// Packs a plan into s.prog (grown when too small) and records the layout in fp.
template <class T>
int upload_plan_t(State& s, const FusedPlan& plan, FusedProgram* fp) {
  std::vector<POp> pops;
  std::vector<FPhase> phases;
  std::vector<rq_c64> poolx;
  if (int rc = pack_ops(plan, sizeof(T) == 4 ? fusion_config().mf : -1, pops, phases, poolx)) return rc;
  const size_t nst = plan.stages.size(), nph = phases.size(), nop = pops.size(), npool = poolx.size();
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_st = 0, o_ph = al(nst * sizeof(FStage)), o_op = o_ph + al(nph * sizeof(FPhase)),
               o_pool = o_op + al(nop * sizeof(POp)), total = o_pool + al(npool * sizeof(Cx<T>));
  if (total > s.prog_cap) {
    if (s.prog) {
      RQ_HIP(hipStreamSynchronize(s.stream));
      RQ_HIP(hipFree(s.prog));
      s.prog = nullptr;
      s.prog_cap = 0;
    }
    size_t cap = total < (1u << 20) ? (1u << 20) : total * 2;
    RQ_HIP(hipMalloc(&s.prog, cap));
    s.prog_cap = cap;
  }
  std::vector<unsigned char> host(total, 0);
  std::memcpy(host.data() + o_st, plan.stages.data(), nst * sizeof(FStage));
  for (FPhase& f : phases)  // the register parts' LDS offsets for this element size
    for (int k = 0; k < (1 << FUSE_R_BITS); ++k) {
      uint32_t t = 0;
      for (int b = 0; b < FUSE_R_BITS; ++b)
        if ((k >> b) & 1) t |= 1u << f.regt[b];
      f.lreg[k] = lds_addr<T>(t);
    }
  std::memcpy(host.data() + o_ph, phases.data(), nph * sizeof(FPhase));
  if constexpr (sizeof(T) == 4) {  // factored 2x2 records carry their f32 coefficients inline
    for (POp& p : pops)
      if ((p.x & 7u) == (uint32_t)POP_U1F) {
        for (int k = 0; k < 2; ++k) {
          p.c[2 * k] = (float)poolx[p.y + k].re;
          p.c[2 * k + 1] = (float)poolx[p.y + k].im;
        }
      } else if ((p.x & 7u) == (uint32_t)FOP_CX && ((p.x >> 5) & 1u)) {  // the relative phase of a CNOT
        p.c[0] = (float)poolx[p.y].re;
        p.c[1] = (float)poolx[p.y].im;
      }
  }
  std::memcpy(host.data() + o_op, pops.data(), nop * sizeof(POp));
  Cx<T>* pp = reinterpret_cast<Cx<T>*>(host.data() + o_pool);
  for (size_t k = 0; k < npool; ++k) pp[k] = Cx<T>{(T)poolx[k].re, (T)poolx[k].im};
  // The host buffer is pageable: wait for the copy before it goes out of scope.
  RQ_HIP(hipMemcpyAsync(s.prog, host.data(), total, hipMemcpyHostToDevice, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  fp->stages = plan.stages;
  fp->o_st = o_st;
  fp->o_ph = o_ph;
  fp->o_op = o_op;
  fp->o_pool = o_pool;
  fp->unitary = plan.unitary;
  return RQ_OK;
}

// One launch per stage of the program in s.prog (asynchronous on s.stream).
template <class T>
int launch_program_t(State& s, const FusedProgram& fp, bool zero_in) {
  const unsigned char* base = reinterpret_cast<const unsigned char*>(s.prog);
  for (size_t i = 0; i < fp.stages.size(); ++i) {
    const FStage& st = fp.stages[i];
    const unsigned threads = 1u << (st.m - FUSE_R_BITS);
    const uint64_t tiles = 1ull << st.nouter;
    hipLaunchKernelGGL(k_fused_stage<T>, dim3((unsigned)tiles), dim3(threads), 0, s.stream, (T*)s.amps,
                       reinterpret_cast<const FStage*>(base + fp.o_st + i * sizeof(FStage)),
                       reinterpret_cast<const FPhase*>(base + fp.o_ph), reinterpret_cast<const POp*>(base + fp.o_op),
                       reinterpret_cast<const Cx<T>*>(base + fp.o_pool), (zero_in && i == 0) ? 1 : 0);
  }
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

int check_shapes(const FusedPlan& plan) {
  for (const FStage& st : plan.stages)
    if (st.m < FUSE_R_BITS || st.m > FUSE_MAX_M || st.nouter > 31)
      return fail(RQ_ERR_ARGUMENT, "fused stage shape out of range");
  return RQ_OK;
}

bool same_config(const FusionConfig& a, const FusionConfig& b) {
  return a.m == b.m && a.c == b.c && a.mf == b.mf && a.seeds == b.seeds;
}

// ---- host execution of the packed program -----------------------------------

// One packed record on host amplitudes, with apply_one's semantics.
template <class T>
void host_apply_one(Cx<T> (&x)[FUSE_R], uint32_t w, const Cx<T>* mp, uint32_t tid, uint64_t tile) {
  const int kind = w & 7u, t0 = (w >> 3) & 3u, t1 = (w >> 5) & 3u, cl = (w >> 7) & 3u, cp = (w >> 9) & 31u,
            dl0 = (w >> 14) & 3u, dp0 = (w >> 16) & 63u, dl1 = (w >> 22) & 3u, dp1 = (w >> 24) & 63u;
  auto mul = [](Cx<T> m, Cx<T> a) { return Cx<T>{m.re * a.re - m.im * a.im, m.re * a.im + m.im * a.re}; };
  auto add = [](Cx<T> a, Cx<T> b) { return Cx<T>{a.re + b.re, a.im + b.im}; };
  auto lv = [&](int loc, int pos) -> int {
    return loc == PL_THR ? (int)((tid >> pos) & 1u) : (int)((tile >> pos) & 1ull);
  };
  auto bit = [&](int loc, int pos, int k) -> int { return loc == PL_REG ? (k >> pos) & 1 : lv(loc, pos); };
  Cx<T> y[FUSE_R];
  for (int k = 0; k < FUSE_R; ++k) y[k] = x[k];
  if (kind == FOP_U1) {
    const Cx<T>* m = (cl == PL_OUT && lv(cl, cp)) ? mp + 4 : mp;
    const int bb = 1 << t0;
    for (int k = 0; k < FUSE_R; ++k) {
      if (k & bb) continue;
      y[k] = add(mul(m[0], x[k]), mul(m[1], x[k | bb]));
      y[k | bb] = add(mul(m[2], x[k]), mul(m[3], x[k | bb]));
    }
  } else if (kind == FOP_U2) {
    const int b0 = 1 << t0, b1 = 1 << t1;
    for (int k = 0; k < FUSE_R; ++k) {
      if (k & (b0 | b1)) continue;
      const int idx[4] = {k, k | b1, k | b0, k | b0 | b1};
      for (int r = 0; r < 4; ++r) {
        Cx<T> acc{0, 0};
        for (int c = 0; c < 4; ++c) acc = add(acc, mul(mp[r * 4 + c], x[idx[c]]));
        y[idx[r]] = acc;
      }
    }
  } else if (kind == FOP_CX) {
    const int bb = 1 << t0;
    Cx<T> xp[FUSE_R];
    for (int k = 0; k < FUSE_R; ++k) xp[k] = ((t1 & 1) && (k & bb)) ? mul(mp[0], x[k]) : x[k];
    for (int k = 0; k < FUSE_R; ++k) {
      const bool on = cl == PL_NONE || (cl == PL_REG ? ((k >> cp) & 1) : lv(cl, cp));
      y[k] = on ? xp[k ^ bb] : xp[k];
    }
  } else if (kind == FOP_CZ) {
    for (int k = 0; k < FUSE_R; ++k)
      if (bit(dl0, dp0, k) & bit(dl1, dp1, k)) y[k] = Cx<T>{-x[k].re, -x[k].im};
  } else if (kind == FOP_D1) {
    for (int k = 0; k < FUSE_R; ++k) y[k] = mul(dl0 == PL_NONE ? mp[k] : mp[bit(dl0, dp0, k)], x[k]);
  } else if (kind == POP_U1F) {
    const int bb = 1 << t0;
    const int anti = t1 ^ (cl == PL_OUT ? lv(cl, cp) : 0);
    for (int k = 0; k < FUSE_R; ++k) {
      if (k & bb) continue;
      if (anti) {
        y[k] = add(mul(mp[0], x[k]), x[k | bb]);
        y[k | bb] = add(x[k], mul(mp[1], x[k | bb]));
      } else {
        y[k] = add(x[k], mul(mp[0], x[k | bb]));
        y[k | bb] = add(mul(mp[1], x[k]), x[k | bb]);
      }
    }
  } else if (kind == FOP_D2) {
    for (int k = 0; k < FUSE_R; ++k) y[k] = mul(mp[2 * bit(dl0, dp0, k) + bit(dl1, dp1, k)], x[k]);
  }  // POP_NOP: unchanged
  for (int k = 0; k < FUSE_R; ++k) x[k] = y[k];
}

// apply_mf on the 64 lanes of one wave (regs[l] = lane l's amplitudes): the
// 32x32x2 f32 matrix-core product spelled out (A[i][k] from lane i + 32k,
// B[k][j] from lane j + 32k, D[i][j] to lane j + 32 ((i / 4) % 2), entry
// (i % 4) + 4 (i / 8)), accumulated in FP64.
void host_apply_mf(Cx<float> (*regs)[FUSE_R], const float* w, uint32_t sel, uint32_t wave, uint64_t tile) {
  uint32_t vb = 0, cb = 0;
  for (int i = 0; i < 3; ++i) {
    const uint32_t d = (sel >> (8 * i)) & 255u, kind = d >> 6, pos = d & 63u;
    if (kind == MF_OUT) vb |= (uint32_t)((tile >> pos) & 1u) << i;
    else if (kind == MF_WAVE) vb |= (((wave * 64u) >> pos) & 1u) << i;
    else if (kind == MF_CHAIN) cb = 1u << i;
  }
  Cx<float> out[64][FUSE_R];
  for (int c = 0; c < 2; ++c) {
    const float* wc = w + (size_t)(vb | (c ? cb : 0u)) * MF_FLOATS;
    for (int i = 0; i < 32; ++i)
      for (int j = 0; j < 32; ++j) {
        double acc = 0.0;
        for (int s = 0; s < 16; ++s)
          for (int kk = 0; kk < 2; ++kk) {
            const double a = wc[(s >> 2) * 256 + (i + 32 * kk) * 4 + (s & 3)];
            const Cx<float>& xv = regs[j + 32 * kk][8 * c + (s >> 1)];
            acc += a * (double)((s & 1) ? xv.im : xv.re);
          }
        const int lane = j + 32 * ((i >> 2) & 1), e = (i & 3) + 4 * (i >> 3);
        Cx<float>& o = out[lane][8 * c + (e >> 1)];
        (e & 1 ? o.im : o.re) = (float)acc;
      }
  }
  for (int l = 0; l < 64; ++l)
    for (int k = 0; k < FUSE_R; ++k) regs[l][k] = out[l][k];
}

template <class T>
void emulate_packed_t(const FusedPlan& plan, const std::vector<POp>& pops, const std::vector<FPhase>& phases,
                      const std::vector<rq_c64>& poolx, Cx<T>* a) {
  std::vector<Cx<T>> pool(poolx.size());
  for (size_t k = 0; k < pool.size(); ++k) pool[k] = Cx<T>{(T)poolx[k].re, (T)poolx[k].im};
  for (const FStage& st : plan.stages) {
    const int m = st.m, nt = m - FUSE_R_BITS;
    const uint32_t nthr = 1u << nt;
    std::vector<Cx<T>> tile((size_t)1 << m);
    std::vector<Cx<T>> regs((size_t)nthr * FUSE_R);
    std::vector<uint32_t> where((size_t)nthr * FUSE_R);
    for (uint64_t tl = 0; tl < (1ull << st.nouter); ++tl) {
      uint64_t outer = 0;
      for (int j = 0; j < st.nouter; ++j) outer |= ((tl >> j) & 1ull) << st.outer[j];
      auto gidx = [&](uint64_t t) {
        uint64_t g = outer;
        for (int b = 0; b < m; ++b) g |= ((t >> b) & 1ull) << st.S[b];
        return g;
      };
      for (uint64_t t = 0; t < tile.size(); ++t) tile[t] = a[gidx(t)];
      for (int p = 0; p < st.nphase; ++p) {
        const FPhase& ph = phases[st.phase0 + p];
        for (uint32_t tid = 0; tid < nthr; ++tid)
          for (int k = 0; k < FUSE_R; ++k) {
            uint32_t t = 0;
            for (int j = 0; j < nt; ++j) t |= ((tid >> j) & 1u) << ph.thrt[j];
            for (int b = 0; b < FUSE_R_BITS; ++b) t |= (uint32_t)((k >> b) & 1) << ph.regt[b];
            where[(size_t)tid * FUSE_R + k] = t;
            regs[(size_t)tid * FUSE_R + k] = tile[t];
          }
        if constexpr (sizeof(T) == 4) {
          if (ph.mf_off >= 0) {
            const float* w = reinterpret_cast<const float*>(pool.data() + ph.mf_off);
            for (uint32_t wv = 0; wv < nthr / 64; ++wv)
              host_apply_mf(reinterpret_cast<Cx<float>(*)[FUSE_R]>(regs.data() + (size_t)wv * 64 * FUSE_R), w,
                            ph.mf_sel, wv, tl);
          }
        }
        for (uint32_t tid = 0; tid < nthr; ++tid) {
          Cx<T>(&x)[FUSE_R] = *reinterpret_cast<Cx<T>(*)[FUSE_R]>(regs.data() + (size_t)tid * FUSE_R);
          for (int o = ph.op_begin; o < ph.op_end; ++o) host_apply_one(x, pops[o].x, pool.data() + pops[o].y, tid, tl);
        }
        for (size_t e = 0; e < regs.size(); ++e) tile[where[e]] = regs[e];
      }
      for (uint64_t t = 0; t < tile.size(); ++t) a[gidx(t)] = tile[t];
    }
  }
}

}  // namespace

int emulate_packed(const FusedPlan& plan, int precision, int mf_min, void* amps) {
  std::vector<POp> pops;
  std::vector<FPhase> phases;
  std::vector<rq_c64> poolx;
  if (int rc = pack_ops(plan, precision == RQ_C64 ? -1 : mf_min, pops, phases, poolx)) return rc;
  if (precision == RQ_C64) emulate_packed_t<double>(plan, pops, phases, poolx, static_cast<Cx<double>*>(amps));
  else emulate_packed_t<float>(plan, pops, phases, poolx, static_cast<Cx<float>*>(amps));
  return RQ_OK;
}

int packed_info(const FusedPlan& plan, int precision, int mf_min, uint64_t* records, uint64_t* mf_phases,
                uint64_t* mf_unitaries) {
  std::vector<POp> pops;
  std::vector<FPhase> phases;
  std::vector<rq_c64> poolx;
  if (int rc = pack_ops(plan, precision == RQ_C64 ? -1 : mf_min, pops, phases, poolx)) return rc;
  uint64_t n = 0, u = 0;
  for (const FPhase& ph : phases) {
    if (ph.mf_off < 0) continue;
    ++n;
    int bits = 0;
    for (int i = 0; i < 3; ++i) bits += ((ph.mf_sel >> (8 * i)) & 255u) != 0;
    u += 1ull << bits;
  }
  if (records) *records = pops.size();
  if (mf_phases) *mf_phases = n;
  if (mf_unitaries) *mf_unitaries = u;
  return RQ_OK;
}

int packed_kinds(const FusedPlan& plan, int precision, int mf_min, uint64_t* counts) {
  std::vector<POp> pops;
  std::vector<FPhase> phases;
  std::vector<rq_c64> poolx;
  if (int rc = pack_ops(plan, precision == RQ_C64 ? -1 : mf_min, pops, phases, poolx)) return rc;
  for (int k = 0; k < 9; ++k) counts[k] = 0;
  for (const POp& p : pops) {
    const uint32_t kind = p.x & 7u, dl0 = (p.x >> 14) & 3u;
    ++counts[(kind == (uint32_t)FOP_D1 && dl0 == (uint32_t)PL_NONE) ? 8 : kind];
  }
  return RQ_OK;
}

int packed_records(const FusedPlan& plan, int precision, int mf_min, uint32_t* words, int32_t* phase, size_t cap,
                   size_t* count) {
  std::vector<POp> pops;
  std::vector<FPhase> phases;
  std::vector<rq_c64> poolx;
  if (int rc = pack_ops(plan, precision == RQ_C64 ? -1 : mf_min, pops, phases, poolx)) return rc;
  *count = pops.size();
  if (pops.size() > cap) return fail(RQ_ERR_ARGUMENT, "record buffer too small");
  for (size_t i = 0; i < pops.size(); ++i) {
    words[i] = pops[i].x;
    phase[i] = -1;
  }
  for (size_t k = 0; k < phases.size(); ++k)
    if (phases[k].mf_off < 0)
      for (int32_t o = phases[k].op_begin; o < phases[k].op_end; ++o) phase[o] = (int32_t)k;
  return RQ_OK;
}

int launch_fused_plan(State& s, const FusedPlan& plan) {
  if (plan.stages.empty()) return RQ_OK;
  if (int rc = check_shapes(plan)) return rc;
  invalidate_fused_program(s);  // s.prog is about to hold another program
  FusedProgram fp;
  int rc = s.prec == RQ_C64 ? upload_plan_t<double>(s, plan, &fp) : upload_plan_t<float>(s, plan, &fp);
  if (rc) return rc;
  return s.prec == RQ_C64 ? launch_program_t<double>(s, fp, false) : launch_program_t<float>(s, fp, false);
}

// Plans (unless the state's cached program was planned from the same gates,
// size, precision and knobs), uploads, and launches one fused segment.
int run_fused_segment(State& s, const rq_gate* gates, size_t ngates, uint64_t* unitary, uint64_t* stages,
                      bool* zero_in) {
  const FusionConfig cfg = fusion_config();
  FusedProgram* fp = s.fused;
  const bool hit = fp && fp->valid && fp->nq == s.nq && fp->prec == s.prec && same_config(fp->cfg, cfg) &&
                   fp->gates.size() == ngates && std::memcmp(fp->gates.data(), gates, ngates * sizeof(rq_gate)) == 0;
  if (!hit) {
    FusedPlan plan;
    int rc = plan_fused(s.nq, gates, ngates, cfg, &plan);
    if (rc) return rc;
    if ((rc = check_shapes(plan))) return rc;
    if (!fp) fp = s.fused = new FusedProgram();
    fp->valid = false;
    rc = s.prec == RQ_C64 ? upload_plan_t<double>(s, plan, fp) : upload_plan_t<float>(s, plan, fp);
    if (rc) return rc;
    fp->gates.assign(gates, gates + ngates);
    fp->nq = s.nq;
    fp->prec = s.prec;
    fp->cfg = cfg;
    fp->valid = true;
  }
  const bool want = zero_in && *zero_in && !fp->stages.empty();
  const bool z = want && s.prec == RQ_C32;
  if (want && !z) launch_init_zero(s);  // complex f64: cleared by a pass of its own
  int rc = s.prec == RQ_C64 ? launch_program_t<double>(s, *fp, z) : launch_program_t<float>(s, *fp, z);
  if (rc) return rc;
  if (want) *zero_in = false;
  *unitary = fp->unitary;
  *stages = fp->stages.size();
  return RQ_OK;
}

void invalidate_fused_program(State& s) {
  if (s.fused) s.fused->valid = false;
}

void release_fused_program(State& s) {
  delete s.fused;
  s.fused = nullptr;
}

}  // namespace rq
