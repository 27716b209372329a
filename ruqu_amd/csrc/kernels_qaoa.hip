// QAOA MaxCut kernels (the state-vector side of qaoa.rs).
//
// * k_fill_uniform: |+>^n as the reference builds it (H on every qubit of
//   |0..0>): every amplitude equals h^n, multiplied up one H at a time.
// * k_cost_phase: the phase separator of one layer in ONE pass.  The reference
//   applies Rzz(2*gamma*w) per edge (qaoa.rs:142-145) as a dense 4x4 diagonal
//   (state.rs:238-276), i.e. a_i <- d_e(bits) * a_i edge after edge; this kernel
//   multiplies every amplitude by the same factors in the same edge order with
//   the same complex product and no contraction, so the result is
//   bit-identical, with one HBM read+write instead of one per edge.
// * k_energy: sum_i |a_i|^2 * C(i), C(i) = sum of weights of the cut edges of
//   bitstring i, in FP64 (one pass; equal to sum_e w*(1-<Z_iZ_j>)/2).
// * k_zmask_batch: <prod Z> for up to 32 Z-only masks in one pass (the
//   per-edge <Z_i Z_j> terms of evaluate_qaoa_cost, qaoa.rs:193-199).
// * k_argmax: index of the largest |a_i|^2, ties to the larger index
//   (Iterator::max_by keeps the last maximum, qaoa.rs:249-255).
#include "qaoa.h"

#pragma clang fp contract(off)

namespace rq {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVE = 64;
constexpr int NB_RED = 2048;

template <class T>
struct Cx {
  T re, im;
};
template <class T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> m, Cx<T> a) {
  return Cx<T>{m.re * a.re - m.im * a.im, m.re * a.im + m.im * a.re};
}

template <class V, class T>
__global__ __launch_bounds__(BLOCK) void k_fill_uniform(V* __restrict__ a, uint64_t n, T c) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) {
    V v;
    v.x = c;
    v.y = 0;
    a[i] = v;
  }
}

// Edge record in device memory: endpoints and the two diagonal factors
// en = e^{-i theta/2} (bits equal) and ep = e^{+i theta/2} (bits differ).
template <class T>
struct EdgeF {
  uint32_t i, j;
  Cx<T> en, ep;
};

template <class V, class T>
__global__ __launch_bounds__(BLOCK) void k_cost_phase(V* __restrict__ a, uint64_t n, const EdgeF<T>* __restrict__ edges,
                                                      int ne) {
  extern __shared__ unsigned char smem[];
  EdgeF<T>* e = reinterpret_cast<EdgeF<T>*>(smem);
  for (int k = threadIdx.x; k < ne; k += BLOCK) e[k] = edges[k];
  __syncthreads();
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i >= n) return;
  V x = a[i];
  Cx<T> v{x.x, x.y};
  for (int k = 0; k < ne; ++k) {
    const uint32_t diff = (uint32_t)(((i >> e[k].i) ^ (i >> e[k].j)) & 1ull);
    v = cmul(diff ? e[k].ep : e[k].en, v);
  }
  x.x = v.re;
  x.y = v.im;
  a[i] = x;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
  return v;
}
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < BLOCK / WAVE; ++k) r += sh[k];
  return r;
}

struct EdgeW {
  uint32_t i, j;
  double w;
};

template <class V>
__global__ __launch_bounds__(BLOCK) void k_energy(const V* __restrict__ a, uint64_t n, const EdgeW* __restrict__ edges,
                                                  int ne, double* __restrict__ part) {
  extern __shared__ unsigned char smem[];
  EdgeW* e = reinterpret_cast<EdgeW*>(smem);
  __shared__ double sh[BLOCK / WAVE];
  for (int k = threadIdx.x; k < ne; k += BLOCK) e[k] = edges[k];
  __syncthreads();
  double acc = 0.0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    V x = a[i];
    double p = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
    double c = 0.0;
    for (int k = 0; k < ne; ++k)
      if (((i >> e[k].i) ^ (i >> e[k].j)) & 1ull) c += e[k].w;
    acc += p * c;
  }
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

constexpr int ZB = 32;
struct ZMasks {
  uint64_t m[ZB];
};

template <class V>
__global__ __launch_bounds__(BLOCK) void k_zmask_batch(const V* __restrict__ a, uint64_t n, ZMasks z, int nz,
                                                       double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  double acc[ZB];
#pragma unroll
  for (int t = 0; t < ZB; ++t) acc[t] = 0.0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    V x = a[i];
    double p = (double)x.x * (double)x.x + (double)x.y * (double)x.y;
#pragma unroll
    for (int t = 0; t < ZB; ++t) acc[t] += (__popcll(i & z.m[t]) & 1) ? -p : p;
  }
#pragma unroll
  for (int t = 0; t < ZB; ++t) {
    if (t < nz) {  // uniform
      double r = block_sum(acc[t], sh);
      if (threadIdx.x == 0) part[(size_t)t * gridDim.x + blockIdx.x] = r;
    }
  }
}

__global__ __launch_bounds__(BLOCK) void k_fold_rows(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  __shared__ double sh[BLOCK / WAVE];
  const double* p = part + (size_t)blockIdx.x * nparts;
  double acc = 0.0;
  for (int k = threadIdx.x; k < nparts; k += BLOCK) acc += p[k];
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

// argmax of |a_i|^2 with ties to the larger index; probabilities computed as
// probabilities() does (f64: re*re+im*im; f32: in f32, then widened).
template <class V>
__device__ __forceinline__ double prob_of(V x);
template <>
__device__ __forceinline__ double prob_of<double2>(double2 x) { return x.x * x.x + x.y * x.y; }
template <>
__device__ __forceinline__ double prob_of<float2>(float2 x) { return (double)(x.x * x.x + x.y * x.y); }

__device__ __forceinline__ void better(double& bv, uint64_t& bi, double v, uint64_t i) {
  // NaN-free probabilities: larger value wins, equal values keep the larger index
  if (v > bv || (v == bv && i > bi)) {
    bv = v;
    bi = i;
  }
}

template <class V>
__global__ __launch_bounds__(BLOCK) void k_argmax(const V* __restrict__ a, uint64_t n, double* __restrict__ pv,
                                                  uint64_t* __restrict__ pi) {
  __shared__ double sv[BLOCK];
  __shared__ uint64_t si[BLOCK];
  double bv = -1.0;
  uint64_t bi = 0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
    better(bv, bi, prob_of(a[i]), i);
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double v = sv[threadIdx.x + s];
      uint64_t i = si[threadIdx.x + s];
      double b = sv[threadIdx.x];
      uint64_t bj = si[threadIdx.x];
      better(b, bj, v, i);
      sv[threadIdx.x] = b;
      si[threadIdx.x] = bj;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = sv[0];
    pi[blockIdx.x] = si[0];
  }
}

__global__ __launch_bounds__(BLOCK) void k_argmax_fold(double* __restrict__ pv, uint64_t* __restrict__ pi, int np) {
  __shared__ double sv[BLOCK];
  __shared__ uint64_t si[BLOCK];
  double bv = -1.0;
  uint64_t bi = 0;
  for (int k = threadIdx.x; k < np; k += BLOCK) better(bv, bi, pv[k], pi[k]);
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = BLOCK / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double b = sv[threadIdx.x];
      uint64_t bj = si[threadIdx.x];
      better(b, bj, sv[threadIdx.x + s], si[threadIdx.x + s]);
      sv[threadIdx.x] = b;
      si[threadIdx.x] = bj;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[0] = sv[0];
    pi[0] = si[0];
  }
}

inline dim3 grid_for(uint64_t items) { return grid_blocks((items + BLOCK - 1) / BLOCK); }
inline unsigned red_blocks(uint64_t n) {
  uint64_t b = (n + BLOCK * 8 - 1) / (BLOCK * 8);
  if (b < 1) b = 1;
  if (b > NB_RED) b = NB_RED;
  return (unsigned)b;
}

}  // namespace

int qaoa_fill_uniform(State& s, double c) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_fill_uniform<double2, double>), grid_for(s.n), dim3(BLOCK), 0, s.stream, (double2*)s.amps,
                       s.n, c);
  else
    hipLaunchKernelGGL((k_fill_uniform<float2, float>), grid_for(s.n), dim3(BLOCK), 0, s.stream, (float2*)s.amps, s.n,
                       (float)c);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

int qaoa_cost_phase(State& s, const void* d_edges, int ne) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL((k_cost_phase<double2, double>), grid_for(s.n), dim3(BLOCK), ne * sizeof(EdgeF<double>),
                       s.stream, (double2*)s.amps, s.n, (const EdgeF<double>*)d_edges, ne);
  else
    hipLaunchKernelGGL((k_cost_phase<float2, float>), grid_for(s.n), dim3(BLOCK), ne * sizeof(EdgeF<float>), s.stream,
                       (float2*)s.amps, s.n, (const EdgeF<float>*)d_edges, ne);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

size_t qaoa_edge_record_bytes(int prec) { return prec == RQ_C64 ? sizeof(EdgeF<double>) : sizeof(EdgeF<float>); }

void qaoa_pack_edges(int prec, const uint32_t* ei, const uint32_t* ej, const rq_c64* en, const rq_c64* ep, int ne,
                     void* out) {
  if (prec == RQ_C64) {
    EdgeF<double>* e = (EdgeF<double>*)out;
    for (int k = 0; k < ne; ++k)
      e[k] = EdgeF<double>{ei[k], ej[k], Cx<double>{en[k].re, en[k].im}, Cx<double>{ep[k].re, ep[k].im}};
  } else {
    // mixed_precision.rs:564-591: f64 matrix entries narrowed to f32
    EdgeF<float>* e = (EdgeF<float>*)out;
    for (int k = 0; k < ne; ++k)
      e[k] = EdgeF<float>{ei[k], ej[k], Cx<float>{(float)en[k].re, (float)en[k].im},
                          Cx<float>{(float)ep[k].re, (float)ep[k].im}};
  }
}

size_t qaoa_weight_record_bytes() { return sizeof(EdgeW); }
void qaoa_pack_weights(const uint32_t* ei, const uint32_t* ej, const double* w, int ne, void* out) {
  EdgeW* e = (EdgeW*)out;
  for (int k = 0; k < ne; ++k) e[k] = EdgeW{ei[k], ej[k], w[k]};
}

int qaoa_energy(State& s, const void* d_wedges, int ne, double* out) {
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, nb + 1);
  if (rc) return rc;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_energy<double2>, dim3(nb), dim3(BLOCK), ne * sizeof(EdgeW), s.stream,
                       (const double2*)s.amps, s.n, (const EdgeW*)d_wedges, ne, s.red + 1);
  else
    hipLaunchKernelGGL(k_energy<float2>, dim3(nb), dim3(BLOCK), ne * sizeof(EdgeW), s.stream, (const float2*)s.amps,
                       s.n, (const EdgeW*)d_wedges, ne, s.red + 1);
  hipLaunchKernelGGL(k_fold_rows, dim3(1), dim3(BLOCK), 0, s.stream, s.red + 1, (int)nb, s.red);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipMemcpyAsync(s.host_red, s.red, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  *out = s.host_red[0];
  return RQ_OK;
}

int reduce_zmasks(State& s, const uint64_t* masks, int nmasks, double* out) {
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, (size_t)ZB * (nb + 1));
  if (rc) return rc;
  for (int t0 = 0; t0 < nmasks; t0 += ZB) {
    int cnt = nmasks - t0 < ZB ? nmasks - t0 : ZB;
    ZMasks z{};
    for (int k = 0; k < cnt; ++k) z.m[k] = masks[t0 + k];
    double* part = s.red + ZB;
    if (s.prec == RQ_C64)
      hipLaunchKernelGGL(k_zmask_batch<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, z,
                         cnt, part);
    else
      hipLaunchKernelGGL(k_zmask_batch<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, z,
                         cnt, part);
    hipLaunchKernelGGL(k_fold_rows, dim3(cnt), dim3(BLOCK), 0, s.stream, part, (int)nb, s.red);
    RQ_HIP(hipGetLastError());
    RQ_HIP(hipMemcpyAsync(s.host_red, s.red, cnt * sizeof(double), hipMemcpyDeviceToHost, s.stream));
    RQ_HIP(hipStreamSynchronize(s.stream));
    for (int k = 0; k < cnt; ++k) out[t0 + k] = s.host_red[k];
  }
  return RQ_OK;
}

int argmax_probability(State& s, uint64_t* index, double* prob) {
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, 2 * (size_t)nb + 2);
  if (rc) return rc;
  double* pv = s.red;
  uint64_t* pi = reinterpret_cast<uint64_t*>(s.red + nb);
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_argmax<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, pv, pi);
  else
    hipLaunchKernelGGL(k_argmax<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, pv, pi);
  hipLaunchKernelGGL(k_argmax_fold, dim3(1), dim3(BLOCK), 0, s.stream, pv, pi, (int)nb);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipMemcpyAsync(s.host_red, pv, sizeof(double), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipMemcpyAsync(s.host_red + 1, pi, sizeof(uint64_t), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  if (prob) *prob = s.host_red[0];
  uint64_t ix;
  std::memcpy(&ix, s.host_red + 1, sizeof(ix));
  if (index) *index = ix;
  return RQ_OK;
}

}  // namespace rq
