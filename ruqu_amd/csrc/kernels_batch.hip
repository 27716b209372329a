// Persistent batched evaluation of QAOA MaxCut and Ry-CZ VQE circuits.
//
// One workgroup evaluates one circuit at a time and takes the next item from
// a device-wide counter (items are sorted largest first), so 10- to 22-qubit
// instances of one batch share a single launch without idle tails.
//
// States are complex-f32; every reduction (energy, Pauli expectations,
// probabilities for the argmax) accumulates in FP64.
//   n <= 13: the whole state lives in LDS for the whole circuit.
//   n >  13: the state lives in this workgroup's HBM slot; each layer streams
//            it through LDS twice (tiles of the low 13 qubits, then tiles that
//            hold the high qubits plus the lowest qubits as contiguous runs).
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
constexpr int TQ = BATCH_TILE_QUBITS;
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

__device__ void qaoa_hbm(float2* s, float2* __restrict__ g, const BInst& in, const BItem& it,
                         const BEdge* __restrict__ edges, const double* __restrict__ params,
                         const float* __restrict__ ctab_all, Shared& sh, double* energy, uint32_t* amax) {
  const int n = in.n, p = in.p, h = n - TQ;
  const uint32_t T = 1u << TQ, N = 1u << n;
  const float* ctab = ctab_all + in.ctab_off;
  const float c0 = plus_amp(n);
  for (int l = 0; l < p; ++l) {
    const double gamma = param_of(params, in, it, l), beta = param_of(params, in, it, p + l);
    float c, sn;
    half_angle(beta, &c, &sn);
    // pass A: tiles of the low TQ qubits (contiguous), cost phase + Rx on them
    for (uint32_t t = 0; t < (1u << h); ++t) {
      const uint32_t base = t << TQ;
      for (uint32_t i = threadIdx.x; i < T; i += B) {
        float2 a = l == 0 ? make_float2(c0, 0.0f) : g[base + i];
        s[i] = phase_mul(a, gamma, (double)ctab[base + i], in.wsum);
      }
      __syncthreads();
      for (int q = 0; q < TQ; ++q) lds_rx(s, TQ, q, c, sn);
      for (uint32_t i = threadIdx.x; i < T; i += B) g[base + i] = s[i];
      __syncthreads();
    }
    // pass B: tiles holding qubits TQ..n-1 (LDS bits TQ-h..TQ-1) and the low
    // TQ-h qubits (contiguous runs); the tile id supplies qubits TQ-h..TQ-1.
    const int lowb = TQ - h;
    for (uint32_t t = 0; t < (1u << h); ++t) {
      for (uint32_t k = threadIdx.x; k < T; k += B) {
        uint32_t lo = k & ((1u << lowb) - 1u), hi = k >> lowb;
        s[k] = g[lo | (t << lowb) | (hi << TQ)];
      }
      __syncthreads();
      for (int q = 0; q < h; ++q) lds_rx(s, TQ, lowb + q, c, sn);
      for (uint32_t k = threadIdx.x; k < T; k += B) {
        uint32_t lo = k & ((1u << lowb) - 1u), hi = k >> lowb;
        g[lo | (t << lowb) | (hi << TQ)] = s[k];
      }
      __syncthreads();
    }
  }
  double acc = 0.0, bv = -1.0;
  uint32_t bi = 0;
  for (uint32_t i = threadIdx.x; i < N; i += B) {
    float2 a = g[i];
    double pr = (double)a.x * a.x + (double)a.y * a.y;
    acc += pr * (double)ctab[i];
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

// sum_k c_k <P_k> over a state of 2^n amplitudes at `a` (LDS or HBM).
__device__ double vqe_energy(const float2* a, int n, const BTerm* __restrict__ terms, int nt, Shared& sh) {
  const uint32_t N = 1u << n;
  double energy = 0.0;
  for (int k = 0; k < nt; ++k) {
    const BTerm t = terms[k];
    double acc = 0.0;
    for (uint32_t i = threadIdx.x; i < N; i += B) {
      float2 ai = a[i], aj = a[i ^ (uint32_t)t.x];
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
      acc += tr * (double)ai.x - ti * (double)ai.y;
    }
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
  *energy = vqe_energy(s, n, terms + in.term_off, in.nterms, sh);
}

__device__ void vqe_hbm(float2* s, float2* __restrict__ g, const BInst& in, const BItem& it,
                        const BTerm* __restrict__ terms, const double* __restrict__ params, Shared& sh,
                        double* energy) {
  const int n = in.n, L = in.p, h = n - TQ, lowb = TQ - h;
  const uint32_t T = 1u << TQ;
  for (int l = 0; l < L; ++l) {
    for (uint32_t t = 0; t < (1u << h); ++t) {  // pass A: Ry on qubits 0..TQ-1
      const uint32_t base = t << TQ;
      for (uint32_t i = threadIdx.x; i < T; i += B)
        s[i] = l == 0 ? make_float2(base + i == 0 ? 1.0f : 0.0f, 0.0f) : g[base + i];
      __syncthreads();
      for (int q = 0; q < TQ; ++q) {
        float c, sn;
        half_angle(param_of(params, in, it, l * n + q) / 2.0, &c, &sn);
        lds_ry(s, TQ, q, c, sn);
      }
      for (uint32_t i = threadIdx.x; i < T; i += B) g[base + i] = s[i];
      __syncthreads();
    }
    for (uint32_t t = 0; t < (1u << h); ++t) {  // pass B: Ry on qubits TQ..n-1, then the CZ chain
      for (uint32_t k = threadIdx.x; k < T; k += B) {
        uint32_t lo = k & ((1u << lowb) - 1u), hi = k >> lowb;
        s[k] = g[lo | (t << lowb) | (hi << TQ)];
      }
      __syncthreads();
      for (int q = 0; q < h; ++q) {
        float c, sn;
        half_angle(param_of(params, in, it, l * n + TQ + q) / 2.0, &c, &sn);
        lds_ry(s, TQ, lowb + q, c, sn);
      }
      for (uint32_t k = threadIdx.x; k < T; k += B) {
        uint32_t lo = k & ((1u << lowb) - 1u), hi = k >> lowb;
        uint32_t gi = lo | (t << lowb) | (hi << TQ);
        g[gi] = cz_chain(s[k], gi, n);
      }
      __syncthreads();
    }
  }
  *energy = vqe_energy(g, n, terms + in.term_off, in.nterms, sh);
}

}  // namespace

__global__ __launch_bounds__(B) void k_batch_eval(const BInst* __restrict__ insts, const BItem* __restrict__ items,
                                                  int nitems, unsigned* __restrict__ counter,
                                                  const int32_t* __restrict__ done, const BEdge* __restrict__ edges,
                                                  const BTerm* __restrict__ terms, const double* __restrict__ params,
                                                  const float* __restrict__ ctab, float2* __restrict__ slots,
                                                  uint64_t slot_amps, double* __restrict__ energy,
                                                  uint32_t* __restrict__ amax) {
  extern __shared__ float2 lds_state[];
  __shared__ Shared sh;
  float2* g = slots + (uint64_t)blockIdx.x * slot_amps;
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
    if (in.kind == BK_QAOA) {
      if (in.n <= LQ) qaoa_lds(lds_state, in, it, edges, params, sh, &en, &am);
      else qaoa_hbm(lds_state, g, in, it, edges, params, ctab, sh, &en, &am);
    } else {
      if (in.n <= LQ) vqe_lds(lds_state, in, it, terms, params, sh, &en);
      else vqe_hbm(lds_state, g, in, it, terms, params, sh, &en);
    }
    if (wid == 0) {  // en and am are block-uniform: wave 0 stores them
      energy[it.out] = en;
      amax[it.out] = am;
    }
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
