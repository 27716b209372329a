// Batched evaluation of QAOA MaxCut and Ry-CZ VQE circuits.
//
// States are complex-f32; every reduction (energy, Pauli expectations,
// probabilities for the argmax) accumulates in FP64.
//   n <= 13 ("small"): k_batch_eval, a persistent kernel; one workgroup
//            evaluates one circuit at a time with the whole state in LDS and
//            takes the next work item from a device counter (items sorted
//            largest first, so uneven sizes share one launch without tails).
//   n >  13 ("wide"): every state of the round has its own region in HBM and
//            is spread over many workgroups, one per LDS tile of 2^tq
//            amplitudes.  Each layer is two launches over all tiles of all
//            wide items: k_wide_a (tiles of the low tq qubits: cost phase,
//            rotations on those qubits) and k_wide_b (tiles holding the high
//            n-tq qubits plus the lowest 2tq-n qubits as contiguous runs:
//            rotations on the high qubits, the CZ chain).  k_wide_energy
//            reduces every tile to an FP64 partial and k_wide_fold folds each
//            item's partials in tile order (deterministic).
//
// QAOA (qaoa.rs:125-153 circuit): |+>^n, then per layer the cost phase
// prod_e Rzz(2 gamma w_e) = exp(i gamma (2 C(i) - W)) per amplitude (C(i) the
// cut weight of bitstring i, W the total weight; phase argument in FP64) and
// Rx(2 beta) on every qubit; energy sum_i |a_i|^2 C(i) and the argmax of
// |a_i|^2 (ties to the larger index, qaoa.rs:249-255).
// VQE: Ry(theta_{l,q}) on every qubit, then CZ(q, q+1) for q < n-1, per
// layer; energy sum_k c_k <P_k> (state.rs:362-408 semantics per term).
#include <hip/hip_runtime.h>

#include <cmath>

#include "batch.h"

namespace rq {
namespace {

constexpr int B = BATCH_THREADS;
constexpr int WAVE = 64;
constexpr int LQ = BATCH_LDS_QUBITS;
constexpr double TWO_PI = 6.283185307179586476925286766559;

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

__device__ __forceinline__ float2 phase_mul(float2 a, double gamma, double cut, double wsum) {
  double phi = gamma * (2.0 * cut - wsum);
  phi -= TWO_PI * rint(phi / TWO_PI);
  float sn, cs;
  sincosf((float)phi, &sn, &cs);
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
__device__ void qaoa_lds(float2* s, const BInst& in, const BItem& it, const BEdge* __restrict__ edges,
                         const double* __restrict__ params, Shared& sh, double* energy, uint32_t* amax) {
  const int n = in.n, p = in.p;
  const uint32_t N = 1u << n;
  const BEdge* e = edges + in.term_off;
  const float c0 = plus_amp(n);
  for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = make_float2(c0, 0.0f);
  __syncthreads();
  for (int l = 0; l < p; ++l) {
    const double gamma = param_of(params, in, it, l), beta = param_of(params, in, it, p + l);
    for (uint32_t i = threadIdx.x; i < N; i += B) s[i] = phase_mul(s[i], gamma, cut_of(i, e, in.nterms), in.wsum);
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
    acc += pr * cut_of(i, e, in.nterms);
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
                                                  double* __restrict__ energy, uint32_t* __restrict__ amax) {
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
    if (in.kind == BK_QAOA) qaoa_lds(lds_state, in, it, edges, params, sh, &en, &am);
    else vqe_lds(lds_state, in, it, terms, params, sh, &en);
    if (wid == 0) {  // en and am are block-uniform: wave 0 stores them
      energy[it.out] = en;
      amax[it.out] = am;
    }
  }
}

// ------------------------------------------------------------ wide items ----
// Block b handles tile (b - wtile0[j]) of wide item j = wmap[b]; its state
// starts at states + wstate[j].  All indices come from blockIdx (uniform).

__global__ __launch_bounds__(B) void k_wide_a(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                              const int32_t* __restrict__ wmap, const int32_t* __restrict__ wtile0,
                                              const int64_t* __restrict__ wstate, const int32_t* __restrict__ done,
                                              const double* __restrict__ params, const float* __restrict__ ctab,
                                              float2* __restrict__ states, int tq, int layer) {
  extern __shared__ float2 s[];
  const int j = wmap[blockIdx.x];
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  if (done[it.inst]) return;
  const uint32_t T = 1u << tq, base = (uint32_t)(blockIdx.x - wtile0[j]) << tq;
  float2* __restrict__ g = states + wstate[j];
  if (in.kind == BK_QAOA) {
    const double gamma = param_of(params, in, it, layer), beta = param_of(params, in, it, in.p + layer);
    const float* __restrict__ ct = ctab + in.ctab_off;
    const float c0 = plus_amp(in.n);
    float c, sn;
    half_angle(beta, &c, &sn);
    for (uint32_t i = threadIdx.x; i < T; i += B) {
      const float2 a = layer == 0 ? make_float2(c0, 0.0f) : g[base + i];
      s[i] = phase_mul(a, gamma, (double)ct[base + i], in.wsum);
    }
    __syncthreads();
    for (int q = 0; q < tq; ++q) lds_rx(s, tq, q, c, sn);
  } else {
    for (uint32_t i = threadIdx.x; i < T; i += B)
      s[i] = layer == 0 ? make_float2(base + i == 0 ? 1.0f : 0.0f, 0.0f) : g[base + i];
    __syncthreads();
    for (int q = 0; q < tq; ++q) {
      float c, sn;
      half_angle(param_of(params, in, it, layer * in.n + q) / 2.0, &c, &sn);
      lds_ry(s, tq, q, c, sn);
    }
  }
  for (uint32_t i = threadIdx.x; i < T; i += B) g[base + i] = s[i];
}

__global__ __launch_bounds__(B) void k_wide_b(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                              const int32_t* __restrict__ wmap, const int32_t* __restrict__ wtile0,
                                              const int64_t* __restrict__ wstate, const int32_t* __restrict__ done,
                                              const double* __restrict__ params, float2* __restrict__ states, int tq,
                                              int layer) {
  extern __shared__ float2 s[];
  const int j = wmap[blockIdx.x];
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  if (done[it.inst]) return;
  const int n = in.n, h = n - tq, lowb = tq - h;
  const uint32_t T = 1u << tq, t = (uint32_t)(blockIdx.x - wtile0[j]), lom = (1u << lowb) - 1u;
  float2* __restrict__ g = states + wstate[j];
  for (uint32_t k = threadIdx.x; k < T; k += B) s[k] = g[(k & lom) | (t << lowb) | ((k >> lowb) << tq)];
  __syncthreads();
  if (in.kind == BK_QAOA) {
    float c, sn;
    half_angle(param_of(params, in, it, in.p + layer), &c, &sn);
    for (int q = 0; q < h; ++q) lds_rx(s, tq, lowb + q, c, sn);
    for (uint32_t k = threadIdx.x; k < T; k += B) g[(k & lom) | (t << lowb) | ((k >> lowb) << tq)] = s[k];
  } else {
    for (int q = 0; q < h; ++q) {
      float c, sn;
      half_angle(param_of(params, in, it, layer * n + tq + q) / 2.0, &c, &sn);
      lds_ry(s, tq, lowb + q, c, sn);
    }
    for (uint32_t k = threadIdx.x; k < T; k += B) {
      const uint32_t gi = (k & lom) | (t << lowb) | ((k >> lowb) << tq);
      g[gi] = cz_chain(s[k], gi, n);
    }
  }
}

// Per-tile partial: QAOA sum |a|^2 C(i) and the tile's argmax of |a|^2; VQE
// sum_k c_k Re(conj(a_{i^x_k}) phase_k(i) a_i) over the tile's indices i.
__global__ __launch_bounds__(B) void k_wide_energy(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                                   const int32_t* __restrict__ wmap,
                                                   const int32_t* __restrict__ wtile0,
                                                   const int64_t* __restrict__ wstate,
                                                   const int32_t* __restrict__ done, const BTerm* __restrict__ terms,
                                                   const float* __restrict__ ctab, const float2* __restrict__ states,
                                                   int tq, BPart* __restrict__ parts) {
  __shared__ Shared sh;
  const int j = wmap[blockIdx.x];
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  if (done[it.inst]) return;
  const uint32_t T = 1u << tq, base = (uint32_t)(blockIdx.x - wtile0[j]) << tq;
  const float2* __restrict__ g = states + wstate[j];
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  double acc = 0.0, bv = -1.0;
  uint32_t bi = 0;
  if (in.kind == BK_QAOA) {
    const float* __restrict__ ct = ctab + in.ctab_off;
    for (uint32_t i = threadIdx.x; i < T; i += B) {
      const float2 a = g[base + i];
      acc += ((double)a.x * a.x + (double)a.y * a.y) * (double)ct[base + i];
      better(bv, bi, (double)(a.x * a.x + a.y * a.y), base + i);
    }
    block_argmax(bv, bi, sh.av, sh.ai);
  } else {
    const BTerm* __restrict__ tt = terms + in.term_off;
    for (int k = 0; k < in.nterms; ++k) {
      const BTerm t = tt[k];
      double a_k = 0.0;
      for (uint32_t i = threadIdx.x; i < T; i += B) {
        const uint32_t gi = base + i;
        a_k += pauli_contrib(g[gi], g[gi ^ (uint32_t)t.x], gi, t);
      }
      acc += t.c * a_k;
    }
  }
  const double e = block_sum_all(acc, sh.red);
  if (wid == 0) parts[blockIdx.x] = BPart{e, bv, bi, 0u};
}

// One block per wide item: folds its tiles' partials (fixed assignment of
// tiles to threads and a fixed reduction tree, so the result does not depend
// on scheduling) into energy[out] and amax[out].
__global__ __launch_bounds__(B) void k_wide_fold(const BInst* __restrict__ insts, const BItem* __restrict__ witems,
                                                 const int32_t* __restrict__ wtile0, const int32_t* __restrict__ done,
                                                 const BPart* __restrict__ parts, int tq, double* __restrict__ energy,
                                                 uint32_t* __restrict__ amax) {
  __shared__ Shared sh;
  const int j = blockIdx.x;
  const BItem it = witems[j];
  const BInst in = insts[it.inst];
  const int wid = __builtin_amdgcn_readfirstlane((int)(threadIdx.x / WAVE));
  if (done[it.inst]) {
    if (wid == 0) {
      energy[it.out] = 0.0;
      amax[it.out] = 0;
    }
    return;
  }
  const int ntiles = 1 << (in.n - tq), t0 = wtile0[j];
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

// Cut table of a QAOA instance held in HBM: C(i) for every bitstring.
__global__ __launch_bounds__(B) void k_cut_table(float* __restrict__ ctab, int n, const BEdge* __restrict__ e, int ne) {
  const uint64_t N = 1ull << n;
  for (uint64_t i = (uint64_t)blockIdx.x * B + threadIdx.x; i < N; i += (uint64_t)gridDim.x * B)
    ctab[i] = (float)cut_of(i, e, ne);
}

__global__ void k_zero_u32(unsigned* __restrict__ p, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}

}  // namespace rq
