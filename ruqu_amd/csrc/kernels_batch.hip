// Batched evaluation of QAOA MaxCut and Ry-CZ VQE circuits.
//
// States are complex-f32; every reduction (energy, Pauli expectations,
// probabilities for the argmax) accumulates in FP64.
//   n <= 13 ("small"): k_batch_eval, a persistent kernel; one workgroup
//            evaluates one circuit at a time with the whole state in LDS and
//            takes the next work item from a device counter (items sorted
//            largest first, so uneven sizes share one launch without tails).
//   n >  13 ("wide"): every state of the round has its own region in HBM and
//            is spread over many workgroups, one per tile of 2^tq amplitudes
//            (tq per state: BInst.tq);
//            k_wide_pass makes p + 1 passes for p layers (see below), the
//            last QAOA pass yields per-tile FP64 energy partials, VQE energies
//            come from k_wide_energy, and k_wide_fold folds each item's
//            partials in tile order (deterministic).
//
// QAOA (qaoa.rs:125-153 circuit): |+>^n, then per layer the cost phase
// prod_e Rzz(2 gamma w_e) = exp(i gamma (2 C(i) - W)) per amplitude (C(i) the
// cut weight of bitstring i, W the total weight; phase argument in FP64) and
// Rx(2 beta) on every qubit; energy sum_i |a_i|^2 C(i) and the argmax of
// |a_i|^2 (ties to the larger index, qaoa.rs:249-255).
// VQE: Ry(theta_{l,q}) on every qubit, then CZ(q, q+1) for q < n-1, per
// layer; energy sum_k c_k <P_k> (state.rs:362-408 semantics per term).
//
// Arithmetic: FP contraction is allowed (the batch has no bit-exact reference
// loop; results are held to the FP64 oracle within 1e-5 of the energy scale),
// so a rotation's outputs are one multiply and one FMA each.
#include <hip/hip_runtime.h>

#include <cmath>

#include "batch.h"

#pragma clang fp contract(fast)

namespace rq {
namespace {

constexpr int B = BATCH_THREADS;
constexpr int WAVE = 64;
constexpr int LQ = BATCH_LDS_QUBITS;
constexpr double TWO_PI = 6.283185307179586476925286766559;
constexpr double INV_TWO_PI = 0.15915494309189533576888376337251;

__device__ __forceinline__ uint32_t ins0(uint32_t p, uint32_t q) {
  uint32_t lo = p & ((1u << q) - 1u);
  return ((p >> q) << (q + 1)) | lo;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
  return v;
}

// Block sum, result broadcast to every thread.
__device__ __forceinline__ double block_sum_all(double v, double* sh) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  for (int k = 0; k < B / WAVE; ++k) r += sh[k];
  __syncthreads();
  return r;
}

__device__ __forceinline__ void better(double& bv, uint32_t& bi, double v, uint32_t i) {
  if (v > bv || (v == bv && i > bi)) {
    bv = v;
    bi = i;
  }
}

// Block argmax (ties to the larger index), result broadcast.
__device__ __forceinline__ void block_argmax(double& bv, uint32_t& bi, double* shv, uint32_t* shi) {
  shv[threadIdx.x] = bv;
  shi[threadIdx.x] = bi;
  __syncthreads();
  for (int s = B / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double v = shv[threadIdx.x];
      uint32_t i = shi[threadIdx.x];
      better(v, i, shv[threadIdx.x + s], shi[threadIdx.x + s]);
      shv[threadIdx.x] = v;
      shi[threadIdx.x] = i;
    }
    __syncthreads();
  }
  bv = shv[0];
  bi = shi[0];
  __syncthreads();
}

// Rotation on qubit q of an LDS-resident block of 2^nb amplitudes.
// Rx(2 beta): [[c, -i s], [-i s, c]];  Ry(theta): [[c, -s], [s, c]].
__device__ __forceinline__ void lds_rx(float2* s, int nb, int q, float c, float sn) {
  const uint32_t half = 1u << (nb - 1);
  for (uint32_t t = threadIdx.x; t < half; t += B) {
    uint32_t i = ins0(t, q), j = i | (1u << q);
    float2 a = s[i], b = s[j];
    s[i] = make_float2(c * a.x + sn * b.y, c * a.y - sn * b.x);
    s[j] = make_float2(sn * a.y + c * b.x, -sn * a.x + c * b.y);
  }
  __syncthreads();
}
__device__ __forceinline__ void lds_ry(float2* s, int nb, int q, float c, float sn) {
  const uint32_t half = 1u << (nb - 1);
  for (uint32_t t = threadIdx.x; t < half; t += B) {
    uint32_t i = ins0(t, q), j = i | (1u << q);
    float2 a = s[i], b = s[j];
    s[i] = make_float2(c * a.x - sn * b.x, c * a.y - sn * b.y);
    s[j] = make_float2(sn * a.x + c * b.x, sn * a.y + c * b.y);
  }
  __syncthreads();
}

// Cut weight of bitstring i.
__device__ __forceinline__ double cut_of(uint64_t i, const BEdge* __restrict__ e, int ne) {
  double c = 0.0;
  for (int k = 0; k < ne; ++k)
    if (((i >> e[k].i) ^ (i >> e[k].j)) & 1ull) c += e[k].w;
  return c;
}

// e^{i phi} a, phi = gamma (2 C - W).  The angle is reduced in FP64 to turns,
// r = phi / 2pi - nearest integer (|r| <= 1/2), and evaluated by the
// hardware sine / cosine, which take turns (v_sin_f32, v_cos_f32: a few
// instructions instead of a full-range sincosf and an FP64 division).
__device__ __forceinline__ float2 phase_mul(float2 a, double gamma, double cut, double wsum) {
  double r = gamma * (2.0 * cut - wsum) * INV_TWO_PI;
  r -= rint(r);
  const float rf = (float)r;
  const float sn = __builtin_amdgcn_sinf(rf), cs = __builtin_amdgcn_cosf(rf);
  return make_float2(cs * a.x - sn * a.y, cs * a.y + sn * a.x);
}

__device__ __forceinline__ double param_of(const double* __restrict__ params, const BInst& in, const BItem& it, int k) {
  double v = params[in.param_off + k];
  if (it.variant == 1 && it.param == k) v += 1.57079632679489661923132169163975144;
  if (it.variant == 2 && it.param == k) v -= 1.57079632679489661923132169163975144;
  return v;
}

__device__ __forceinline__ void half_angle(double theta_half, float* c, float* s) {
  double sd, cd;
  sincos(theta_half, &sd, &cd);
  *c = (float)cd;
  *s = (float)sd;
}

__device__ __forceinline__ float plus_amp(int n) {
  const float h = 0.70710678118654752440f;
  float c = 1.0f;
  for (int q = 0; q < n; ++q) c = h * c;
  return c;
}

struct Shared {
  double red[B / WAVE];
  double av[B];
  uint32_t ai[B];
  int item;
};

// ---------------------------------------------------------------- QAOA ----
__device__ void qaoa_lds(float2* s, const BInst& in, const BItem& it, const float* __restrict__ ctab,
                         const double* __restrict__ params, Shared& sh, double* energy, uint32_t* amax) {
  const int n = in.n, p = in.p;
  const uint32_t N = 1u << n;
  const float* __restrict__ ct = ctab + in.ctab_off;  // C(i), from k_cut_table
  const float c0 = plus_amp(n);
  for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = make_float2(c0, 0.0f);
  __syncthreads();
  for (int l = 0; l < p; ++l) {
    const double gamma = param_of(params, in, it, l), beta = param_of(params, in, it, p + l);
    for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = phase_mul(s[i], gamma, (double)ct[i], in.wsum);
    __syncthreads();
    float c, sn;
    half_angle(beta, &c, &sn);  // Rx(2 beta): half angle beta
    for (int q = 0; q < n; ++q) lds_rx(s, n, q, c, sn);
  }
  double acc = 0.0, bv = -1.0;
  uint32_t bi = 0;
  for (uint32_t i = threadIdx.x; i < N; i += B) {
    float2 a = s[i];
    double pr = (double)a.x * a.x + (double)a.y * a.y;
    acc += pr * (double)ct[i];
    better(bv, bi, (double)(a.x * a.x + a.y * a.y), i);
  }
  *energy = block_sum_all(acc, sh.red);
  block_argmax(bv, bi, sh.av, sh.ai);
  *amax = bi;
}

// ----------------------------------------------------------------- VQE ----
__device__ __forceinline__ float2 cz_chain(float2 a, uint32_t i, int n) {
  const uint32_t m = (1u << (n - 1)) - 1u;
  return (__popc(i & (i >> 1) & m) & 1) ? make_float2(-a.x, -a.y) : a;
}

// Re((conj(a_j) * phase) * a_i) for Pauli term t at index i (state.rs:362-400
// semantics: X/Y flip bit q, Z/Y sign from bit q of i, i^{nY} overall).
__device__ __forceinline__ double pauli_contrib(float2 ai, float2 aj, uint32_t i, const BTerm& t) {
  double s = (__popc(i & (uint32_t)t.zy) & 1) ? -1.0 : 1.0;
  double pr, pim;
  switch (t.ny & 3) {
    case 0: pr = s; pim = 0.0; break;
    case 1: pr = 0.0; pim = s; break;
    case 2: pr = -s; pim = 0.0; break;
    default: pr = 0.0; pim = -s; break;
  }
  double cr = aj.x, ci = -(double)aj.y;
  double tr = cr * pr - ci * pim, ti = cr * pim + ci * pr;
  return tr * (double)ai.x - ti * (double)ai.y;
}

// sum_k c_k <P_k> over an LDS-resident state of 2^n amplitudes.
__device__ double vqe_energy_lds(const float2* a, int n, const BTerm* __restrict__ terms, int nt, Shared& sh) {
  const uint32_t N = 1u << n;
  double energy = 0.0;
  for (int k = 0; k < nt; ++k) {
    const BTerm t = terms[k];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < N; i += B) acc += pauli_contrib(a[i], a[i ^ (uint32_t)t.x], i, t);
    energy += t.c * block_sum_all(acc, sh.red);
  }
  return energy;
}

__device__ void vqe_lds(float2* s, const BInst& in, const BItem& it, const BTerm* __restrict__ terms,
                        const double* __restrict__ params, Shared& sh, double* energy) {
  const int n = in.n, L = in.p;
  const uint32_t N = 1u << n;
  for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = make_float2(i == 0 ? 1.0f : 0.0f, 0.0f);
  __syncthreads();
  for (int l = 0; l < L; ++l) {
    for (int q = 0; q < n; ++q) {
      float c, sn;
      half_angle(param_of(params, in, it, l * n + q) / 2.0, &c, &sn);
      lds_ry(s, n, q, c, sn);
    }
    for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = cz_chain(s[i], i, n);
    __syncthreads();
  }
  *energy = vqe_energy_lds(s, n, terms + in.term_off, in.nterms, sh);
}

}  // namespace

__global__ __launch_bounds__(B) void k_batch_eval(const BInst* __restrict__ insts, const BItem* __restrict__ items,
                                                  int nitems, unsigned* __restrict__ counter,
                                                  const int32_t* __restrict__ done, const BEdge* __restrict__ edges,
                                                  const BTerm* __restrict__ terms, const double* __restrict__ params,
                                                  const float* __restrict__ ctab, double* __restrict__ energy,
                                                  uint32_t* __restrict__ amax) {
  extern __shared__ float2 lds_state[];
  __shared__ Shared sh;
  // Every branch of this loop is uniform across a wave (wave index and item
  // index come through readfirstlane).  A branch on the lane index here
  // (`if (threadIdx.x == 0)` around the ticket or the result store) lets the
  // compiler split the loop at that branch into an inner loop, left by
  // thread 0 alone after each item: the other lanes then re-run the same
  // item forever while thread 0 waits to take the next ticket.
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const bool lane0 = (threadIdx.x & (WAVE - 1)) == 0;
  // A block takes at most nitems tickets that are in range, so the trip
  // count is bounded even if tickets stopped advancing.
  for (int taken = 0; taken <= nitems; ++taken) {
    if (wid == 0) {
      // One increment per block (lane 0 adds 1, the others 0); every lane of
      // wave 0 writes lane 0's ticket.
      const int v = __builtin_amdgcn_readfirstlane((int)atomicAdd(counter, lane0 ? 1u : 0u));
      sh.item = v;
    }
    __syncthreads();
    const int k = __builtin_amdgcn_readfirstlane(sh.item);
    __syncthreads();
    if (k >= nitems) break;
    const BItem it = items[k];
    const BInst in = insts[it.inst];
    if (done[it.inst]) {
      if (wid == 0) {
        energy[it.out] = 0.0;
        amax[it.out] = 0;
      }
      continue;
    }
    double en = 0.0;
    uint32_t am = 0;
    if (in.kind == BK_QAOA) qaoa_lds(lds_state, in, it, ctab, params, sh, &en, &am);
    else vqe_lds(lds_state, in, it, terms, params, sh, &en);
    if (wid == 0) {  // en and am are block-uniform: wave 0 stores them
      energy[it.out] = en;
      amax[it.out] = am;
    }
  }
}

// ------------------------------------------------------------ wide items ----
// A wide state of n qubits is cut into tiles of 2^tq amplitudes (h = n - tq
// high qubits, lowb = tq - h) in one of two tilings:
//   A: tile t = global bits tq..n-1; tile bits 0..tq-1 = global bits 0..tq-1.
//   B: tile t = global bits lowb..tq-1; tile bits 0..lowb-1 = global bits
//      0..lowb-1 and tile bits lowb..tq-1 = global bits tq..n-1.
// A tiling rotates its "own" qubits: A the low tq, B the high h.  Pass s uses
// tiling A for even s and B for odd s:
//   QAOA, p layers, passes 0..p: pass s >= 1 first completes layer s-1 (Rx on
//   its own qubits); for s < p it then applies the cost phase of layer s (a
//   diagonal, valid in either tiling) and Rx of layer s on its own qubits.
//   Pass p ends with the tile's energy partial and argmax instead of a store.
//   VQE, L layers, passes 0..L: pass s >= 1 completes layer s-1 (Ry on its
//   own qubits, then the CZ chain); for s < L it then applies Ry of layer s
//   on its own qubits.
// So p layers cost p + 1 passes over the state.  Inside a pass each thread
// holds 2^(tq-8) amplitudes in registers: those whose tile indices differ in
// a window of tq-8 tile bits; rotations on the window's bits are register
// butterflies, and the window moves through an LDS transpose.

// LDS slot of tile index k: one pad slot every 2^r entries (bank spread when
// a thread's registers are contiguous tile indices).
__device__ __forceinline__ uint32_t lds_slot(uint32_t k, int r) { return k + (k >> r); }

// Tile index of register m of thread `tid` for a window starting at bit lo.
template <int R>
__device__ __forceinline__ uint32_t win_index(uint32_t tid, int lo, int m) {
  return (tid & ((1u << lo) - 1u)) | ((uint32_t)m << lo) | ((tid >> lo) << (lo + R));
}

// Workgroups are dispatched to the 8 XCDs round-robin.  Within each chunk of
// 32 workgroups this sends tiles 4x..4x+3 to XCD x, so the tiles of tiling B
// that share 128-byte lines (lowb < 4) are read through one L2.
__device__ __forceinline__ uint32_t xcd_swizzle(uint32_t b, uint32_t grid) {
  constexpr uint32_t G = 4, C = 8 * G;
  if (b >= grid - grid % C) return b;
  const uint32_t r = b % C;
  return b - r + (r % 8) * G + r / 8;
}

template <int TQ>
struct WideCtx {
  static constexpr int R = TQ - 8, NR = 1 << R;
  uint32_t t;
  int n, lowb;
  bool tb;  // tiling B
  __device__ __forceinline__ uint32_t gidx(uint32_t k) const {
    if (!tb) return (t << TQ) | k;
    const uint32_t lom = (1u << lowb) - 1u;
    return (k & lom) | (t << lowb) | ((k >> lowb) << TQ);
  }
  __device__ __forceinline__ int qubit_of(int bit) const { return tb ? TQ + bit - lowb : bit; }
};

// The thread's part of a window's tile indices and m << lo have disjoint
// bits, so the LDS slot of register m is a per-thread base plus a uniform
// offset (one add per access; the offsets are scalar).
template <int R>
__device__ __forceinline__ uint32_t win_slot_base(uint32_t tid, int lo) {
  const uint32_t t = (tid & ((1u << lo) - 1u)) | ((tid >> lo) << (lo + R));
  return t + (t >> R);
}
template <int R>
__device__ __forceinline__ uint32_t win_slot_off(int lo, int m) {
  const uint32_t k = (uint32_t)m << lo;
  return k + (k >> R);
}

template <int TQ>
__device__ __forceinline__ void wide_transpose(float2 (&v)[1 << (TQ - 8)], int from, int to, float2* lds) {
  constexpr int R = TQ - 8, NR = 1 << R;
  const uint32_t bf = win_slot_base<R>(threadIdx.x, from), bt = win_slot_base<R>(threadIdx.x, to);
#pragma unroll
  for (int m = 0; m < NR; ++m) lds[bf + win_slot_off<R>(from, m)] = v[m];
  __syncthreads();
#pragma unroll
  for (int m = 0; m < NR; ++m) v[m] = lds[bt + win_slot_off<R>(to, m)];
  __syncthreads();
}

// A copy of x in new registers (see rot_step).
__device__ __forceinline__ float fresh(float x) {
  float r;
  asm("v_mov_b32 %0, %1" : "=v"(r) : "v"(x));
  return r;
}

// One bit of a window rotation, x -> y (distinct register arrays), with the
// larger of (c, s) = (cos, sin) of half the angle factored out of the 2x2
// (the factor goes into the pass's uniform scale, see rot_factor), so that
// a pair costs 4 FMAs with a ratio k of magnitude <= 1:
//   mode 1: Rx = c [[1, -ik], [-ik, 1]], k = s / c
//   mode 2: Ry = c [[1, -k], [k, 1]],    k = s / c
//   mode 3: Rx = -is [[k, 1], [1, k]] with outputs (b + ika, a + ikb), k = c / s
//   mode 4: Ry = s [[k, -1], [1, k]],    k = c / s
// mode 0: the bit is not rotated, an explicit copy.  Alternating the two
// arrays and copying explicitly keeps every step's results in fresh
// registers; with in-place updates under the uniform branches the compiler
// re-merges the registers after each step and moves half the amplitudes
// every time.
template <int NR, int JB>
__device__ __forceinline__ void rot_step(const float2 (&x)[NR], float2 (&y)[NR], int mode, float k) {
  if (mode == 1) {
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      if (m & (1 << JB)) continue;
      const float2 a = x[m], b = x[m | (1 << JB)];
      y[m] = make_float2(a.x + k * b.y, a.y - k * b.x);
      y[m | (1 << JB)] = make_float2(b.x + k * a.y, b.y - k * a.x);
    }
  } else if (mode == 2) {
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      if (m & (1 << JB)) continue;
      const float2 a = x[m], b = x[m | (1 << JB)];
      y[m] = make_float2(a.x - k * b.x, a.y - k * b.y);
      y[m | (1 << JB)] = make_float2(b.x + k * a.x, b.y + k * a.y);
    }
  } else if (mode == 3) {
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      if (m & (1 << JB)) continue;
      const float2 a = x[m], b = x[m | (1 << JB)];
      y[m] = make_float2(b.x - k * a.y, b.y + k * a.x);
      y[m | (1 << JB)] = make_float2(a.x - k * b.y, a.y + k * b.x);
    }
  } else if (mode == 4) {
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      if (m & (1 << JB)) continue;
      const float2 a = x[m], b = x[m | (1 << JB)];
      y[m] = make_float2(k * a.x - b.x, k * a.y - b.y);
      y[m | (1 << JB)] = make_float2(a.x + k * b.x, a.y + k * b.y);
    }
  } else {
#pragma unroll
    for (int m = 0; m < NR; ++m) y[m] = make_float2(fresh(x[m].x), fresh(x[m].y));
  }
}

// Factored form of a rotation by (c, s) = (cos, sin) of half its angle:
// mode offset (0: c factored out, 2: s), ratio k and the factor f that the
// pass's scale takes (Rx: c or -is; Ry: c or s).
struct RotFactor {
  float k;
  int big_s;
  float2 fx, fy;  // factor of Rx, of Ry
};
__device__ __forceinline__ RotFactor rot_factor(double c, double s) {
  RotFactor r;
  if (fabs(c) >= fabs(s)) {
    r.big_s = 0;
    r.k = (float)(s / c);
    r.fx = make_float2((float)c, 0.0f);
    r.fy = make_float2((float)c, 0.0f);
  } else {
    r.big_s = 1;
    r.k = (float)(c / s);
    r.fx = make_float2(0.0f, (float)-s);
    r.fy = make_float2((float)s, 0.0f);
  }
  return r;
}

// Steps JB .. R-1 of a window rotation, alternating a -> b -> a ...; the
// result ends in v.
template <int TQ, int JB>
__device__ __forceinline__ void rot_chain(float2 (&a)[1 << (TQ - 8)], float2 (&b)[1 << (TQ - 8)],
                                          float2 (&v)[1 << (TQ - 8)], const int* mode, const float* k) {
  constexpr int NR = 1 << (TQ - 8);
  if constexpr (JB < TQ - 8) {
    rot_step<NR, JB>(a, b, mode[JB], k[JB]);
    rot_chain<TQ, JB + 1>(b, a, v, mode, k);
  } else {
#pragma unroll
    for (int m = 0; m < NR; ++m) v[m] = a[m];
  }
}

// Rotations of one layer on the tile bits of window lo that are own bits not
// yet rotated in this layer (done: mask of tile bits).  rx: Rx, else Ry, in
// the factored form of rf[q]; the factors multiply into scale (the same in
// every thread: every amplitude of the state gets the same rotations).
template <int TQ>
__device__ __forceinline__ void wide_rotate(float2 (&v)[1 << (TQ - 8)], int lo, int own_lo, uint32_t& done,
                                            const WideCtx<TQ>& cx, bool rx, const RotFactor* rf, float2& scale) {
  constexpr int R = TQ - 8, NR = 1 << R;
  int mode[R];
  float k[R];
#pragma unroll
  for (int jb = 0; jb < R; ++jb) {
    const int bit = lo + jb;
    mode[jb] = 0;
    k[jb] = 0.0f;
    if (bit < own_lo || ((done >> bit) & 1u)) continue;
    done |= 1u << bit;
    const RotFactor& f = rf[cx.qubit_of(bit)];
    mode[jb] = (rx ? 1 : 2) + 2 * f.big_s;
    k[jb] = f.k;
    const float2 g = rx ? f.fx : f.fy;
    scale = make_float2(scale.x * g.x - scale.y * g.y, scale.x * g.y + scale.y * g.x);
  }
  float2 w[NR];
  rot_chain<TQ, 0>(v, w, v, mode, k);
}

template <int TQ>
__global__ __launch_bounds__(B) void k_wide_pass(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                                 const int32_t* __restrict__ wmap,
                                                 const int32_t* __restrict__ wtile0,
                                                 const int64_t* __restrict__ wstate, const WidePasses wp,
                                                 const int32_t* __restrict__ done_flags,
                                                 const double* __restrict__ params, const float* __restrict__ ctab,
                                                 float2* __restrict__ states, int pass, BPart* __restrict__ parts) {
  constexpr int R = TQ - 8, NR = 1 << R;
  extern __shared__ float2 lds[];  // 2^TQ + 2^(TQ-R) slots
  __shared__ RotFactor rf[2][BATCH_MAX_QUBITS];
  __shared__ Shared sh;
  // this pass's tile list (round tile ids); item j's first pass reads wp.src
  const uint32_t b = (uint32_t)wp.ptile[xcd_swizzle(blockIdx.x, gridDim.x)];
  const int j = wmap[b];
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  if (done_flags[it.inst]) return;
  const bool qaoa = in.kind == BK_QAOA;
  const int L = in.p;
  const bool l1 = pass >= 1, l2 = pass < L;
  WideCtx<TQ> cx;
  cx.n = in.n;
  cx.lowb = 2 * TQ - in.n;
  cx.tb = pass & 1;
  cx.t = b - (uint32_t)wtile0[j];
  const int own_lo = cx.tb ? cx.lowb : 0;
  float2* __restrict__ g = states + wstate[j];
  // rotation tables: [0] layer pass-1, [1] layer pass (factored half angles)
  if (threadIdx.x < 2 * BATCH_MAX_QUBITS) {
    const int sel = threadIdx.x / BATCH_MAX_QUBITS, q = threadIdx.x % BATCH_MAX_QUBITS;
    const int layer = pass - 1 + sel;
    double c = 1.0, sn = 0.0;
    if (q < in.n && layer >= 0 && layer < L) {
      const double h = qaoa ? param_of(params, in, it, L + layer) : param_of(params, in, it, layer * in.n + q) / 2.0;
      sincos(h, &sn, &c);
    }
    rf[sel][q] = rot_factor(c, sn);
  }
  __syncthreads();
  float2 scale = make_float2(1.0f, 0.0f);  // product of the factored-out rotation factors
  const int nw = (TQ - own_lo + R - 1) / R;  // windows covering the own bits, top first
  float2 v[NR];
  int lo;
  if (pass == 0) {
    lo = 0;
    const float c0 = plus_amp(in.n);
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      const uint32_t gi = cx.gidx(win_index<R>(threadIdx.x, lo, m));
      v[m] = qaoa ? make_float2(c0, 0.0f) : make_float2(gi == 0 ? 1.0f : 0.0f, 0.0f);
    }
  } else {
    lo = TQ - R;
    const float2* __restrict__ src = pass == wp.first[j] ? states + wp.src[j] : g;
#pragma unroll
    for (int m = 0; m < NR; ++m) v[m] = src[cx.gidx(win_index<R>(threadIdx.x, lo, m))];
  }
  const float* __restrict__ ct = qaoa ? ctab + in.ctab_off : nullptr;
  if (l1) {  // complete layer pass-1: windows top -> bottom
    uint32_t dm = 0;
    for (int w = 0; w < nw; ++w) {
      const int wl = max(TQ - R * (w + 1), 0);
      if (wl != lo) {
        wide_transpose<TQ>(v, lo, wl, lds);
        lo = wl;
      }
      wide_rotate<TQ>(v, lo, own_lo, dm, cx, qaoa, rf[0], scale);
    }
    if (!qaoa) {
#pragma unroll
      for (int m = 0; m < NR; ++m) v[m] = cz_chain(v[m], cx.gidx(win_index<R>(threadIdx.x, lo, m)), in.n);
    }
  }
  if (l2) {  // layer pass: (QAOA cost phase,) rotations bottom -> top
    if (qaoa) {
      const double gamma = param_of(params, in, it, pass);
#pragma unroll
      for (int m = 0; m < NR; ++m) {
        const uint32_t gi = cx.gidx(win_index<R>(threadIdx.x, lo, m));
        v[m] = phase_mul(v[m], gamma, (double)ct[gi], in.wsum);
      }
    }
    uint32_t dm = 0;
    for (int w = nw - 1; w >= 0; --w) {
      const int wl = max(TQ - R * (w + 1), 0);
      if (wl != lo) {
        wide_transpose<TQ>(v, lo, wl, lds);
        lo = wl;
      }
      wide_rotate<TQ>(v, lo, own_lo, dm, cx, qaoa, rf[1], scale);
    }
  }
  if (!l2 && qaoa) {  // last QAOA pass: energy partial and argmax of the tile
    double acc = 0.0, bv = -1.0;
    uint32_t bi = 0;
#pragma unroll
    for (int m = 0; m < NR; ++m) {
      const uint32_t gi = cx.gidx(win_index<R>(threadIdx.x, lo, m));
      const float2 a = v[m];
      acc += ((double)a.x * a.x + (double)a.y * a.y) * (double)ct[gi];
      better(bv, bi, (double)(a.x * a.x + a.y * a.y), gi);
    }
    // the pass's rotation factors scale every probability by |scale|^2
    const double s2 = (double)scale.x * scale.x + (double)scale.y * scale.y;
    acc *= s2;
    bv *= s2;
    const double e = block_sum_all(acc, sh.red);
    block_argmax(bv, bi, sh.av, sh.ai);
    const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
    if (wid == 0) parts[b] = BPart{e, bv, bi, 0u};
    return;
  }
#pragma unroll
  for (int m = 0; m < NR; ++m) {
    const float2 a = v[m];
    g[cx.gidx(win_index<R>(threadIdx.x, lo, m))] =
        make_float2(scale.x * a.x - scale.y * a.y, scale.x * a.y + scale.y * a.x);
  }
}

template __global__ void k_wide_pass<9>(const BInst*, const BItem*, const int32_t*, const int32_t*, const int64_t*,
                                        const WidePasses, const int32_t*, const double*, const float*, float2*, int, BPart*);
template __global__ void k_wide_pass<10>(const BInst*, const BItem*, const int32_t*, const int32_t*, const int64_t*,
                                         const WidePasses, const int32_t*, const double*, const float*, float2*, int, BPart*);
template __global__ void k_wide_pass<11>(const BInst*, const BItem*, const int32_t*, const int32_t*, const int64_t*,
                                         const WidePasses, const int32_t*, const double*, const float*, float2*, int, BPart*);
template __global__ void k_wide_pass<12>(const BInst*, const BItem*, const int32_t*, const int32_t*, const int64_t*,
                                         const WidePasses, const int32_t*, const double*, const float*, float2*, int, BPart*);
template __global__ void k_wide_pass<13>(const BInst*, const BItem*, const int32_t*, const int32_t*, const int64_t*,
                                         const WidePasses, const int32_t*, const double*, const float*, float2*, int, BPart*);

// VQE tile partial sum_k c_k Re(conj(a_{i^x_k}) phase_k(i) a_i) over the
// tile's indices i (natural order: tile t = indices t*2^tq ..).  Block b
// handles round tile vmap[b].
__global__ __launch_bounds__(B) void k_wide_energy(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                                   const int32_t* __restrict__ wmap,
                                                   const int32_t* __restrict__ wtile0,
                                                   const int64_t* __restrict__ wstate,
                                                   const int32_t* __restrict__ vmap,
                                                   const int32_t* __restrict__ done, const BTerm* __restrict__ terms,
                                                   const float2* __restrict__ states,
                                                   BPart* __restrict__ parts) {
  __shared__ Shared sh;
  const int tile = vmap[blockIdx.x];
  const int j = wmap[tile];
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  if (done[it.inst]) return;
  const uint32_t T = 1u << in.tq, base = (uint32_t)(tile - wtile0[j]) << in.tq;
  const float2* __restrict__ g = states + wstate[j];
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  const BTerm* __restrict__ tt = terms + in.term_off;
  double acc = 0.0;
  for (int k = 0; k < in.nterms; ++k) {
    const BTerm t = tt[k];
    double a_k = 0.0;
    for (uint32_t i = threadIdx.x; i < T; i += B) {
      const uint32_t gi = base + i;
      a_k += pauli_contrib(g[gi], g[gi ^ (uint32_t)t.x], gi, t);
    }
    acc += t.c * a_k;
  }
  const double e = block_sum_all(acc, sh.red);
  if (wid == 0) parts[tile] = BPart{e, -1.0, 0u, 0u};
}

// One block per wide item: folds its tiles' partials (fixed assignment of
// tiles to threads and a fixed reduction tree, so the result does not depend
// on scheduling) into energy[out] and amax[out].
__global__ __launch_bounds__(B) void k_wide_fold(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                                 const int32_t* __restrict__ wtile0, const int32_t* __restrict__ done,
                                                 const BPart* __restrict__ parts, double* __restrict__ energy,
                                                 uint32_t* __restrict__ amax) {
  __shared__ Shared sh;
  const int j = blockIdx.x;
  const BItem it = witems[j];
  if (it.out < 0) return;  // an advance item (shared prefix state): no energy
  const BInst in = insts[it.inst];
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  if (done[it.inst]) {
    if (wid == 0) {
      energy[it.out] = 0.0;
      amax[it.out] = 0;
    }
    return;
  }
  const int ntiles = 1 << (in.n - in.tq), t0 = wtile0[j];
  double acc = 0.0, bv = -1.0;
  uint32_t bi = 0;
  for (int t = threadIdx.x; t < ntiles; t += B) {
    const BPart q = parts[t0 + t];
    acc += q.e;
    better(bv, bi, q.pm, q.idx);
  }
  const double e = block_sum_all(acc, sh.red);
  block_argmax(bv, bi, sh.av, sh.ai);
  if (wid == 0) {
    energy[it.out] = e;
    amax[it.out] = in.kind == BK_QAOA ? bi : 0u;
  }
}

// Cut tables of every QAOA instance of a batch, C(i) for every bitstring,
// in one launch: workgroup w (rows of the grid in blockIdx.y; the last row
// may run past nchunks) fills BATCH_CUT_CHUNK entries of one table.
__global__ __launch_bounds__(B) void k_cut_tables(float* __restrict__ ctab, const BInst* __restrict__ insts,
                                                  const BEdge* __restrict__ edges,
                                                  const BCutChunk* __restrict__ chunks, uint32_t nchunks) {
  const uint64_t w = (uint64_t)blockIdx.y * gridDim.x + blockIdx.x;
  if (w >= nchunks) return;
  const BCutChunk ch = chunks[w];
  const BInst& in = insts[ch.inst];
  const uint64_t N = 1ull << in.n, first = (uint64_t)ch.chunk * BATCH_CUT_CHUNK;
  const uint64_t end = N < first + BATCH_CUT_CHUNK ? N : first + BATCH_CUT_CHUNK;
  float* __restrict__ t = ctab + in.ctab_off;
  const BEdge* __restrict__ e = edges + in.term_off;
  for (uint64_t i = first + threadIdx.x; i < end; i += B) t[i] = (float)cut_of(i, e, in.nterms);
}

// Maximum cut weight over all bitstrings; cut weights are non-negative, so
// their IEEE bit patterns order like their values and the block maxima
// combine with an integer atomic max (result independent of order).
__global__ __launch_bounds__(B) void k_maxcut(int n, const BEdge* __restrict__ e, int ne,
                                              unsigned long long* __restrict__ out) {
  __shared__ double red[B];
  const uint64_t N = 1ull << n;
  double best = 0.0;
  for (uint64_t i = (uint64_t)blockIdx.x * B + threadIdx.x; i < N; i += (uint64_t)gridDim.x * B)
    best = fmax(best, cut_of(i, e, ne));
  red[threadIdx.x] = best;
  __syncthreads();
  for (int s = B / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(out, (unsigned long long)__double_as_longlong(red[0]));
}

__global__ void k_zero_u32(unsigned* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace rq
