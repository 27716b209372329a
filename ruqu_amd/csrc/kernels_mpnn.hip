// Kernels of the warm-start graph network (forward, backward, Adam).
//
// Matrix products run on the CDNA4 matrix cores: v_mfma_f32_32x32x16_bf16,
// bfloat16 operands, FP32 accumulation.  Everything else (message
// aggregation, pooling, activations, loss, optimizer) is FP32 VALU work.
// Rounding to bfloat16 is round-to-nearest-even on the FP32 bits, at the
// points where a value enters a matrix product.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "mpnn.h"

namespace rq {
namespace mpnn {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

__device__ __forceinline__ uint16_t f2bf(float x) {
  uint32_t u = __float_as_uint(x);
  u += 0x7FFFu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

namespace {

constexpr int BM = 64, BN = 64, BK = 32, LDK = BK + 8;

// C[M x N] = A[M x K] B[K x N] (+ bias[n]).  A(m,k) = A_KC ? A[m*lda+k] :
// A[k*lda+m]; B(k,n) = B_KC ? B[n*ldb+k] : B[k*ldb+n] (bfloat16 bits).
// Block tile 64 x 64 (4 waves of 32 x 32), K staged through LDS 32 at a time
// in K-contiguous rows.  M and K may be ragged (zero fill); N is a multiple
// of 64.  Matrix-core operand layout (32x32x16): lane l holds row (A) or
// column (B) l % 32 and the 8 K values of block l / 32; the 16 results of
// lane l are column l % 32, rows 8*(i/4) + 4*(l/32) + i%4.
template <bool A_KC, bool B_KC>
__global__ __launch_bounds__(256) void k_gemm_bf16(int M, int N, int K, const uint16_t* __restrict__ A, int lda,
                                                   const uint16_t* __restrict__ B, int ldb, float* __restrict__ C,
                                                   int ldc, const float* __restrict__ bias) {
  __shared__ __attribute__((aligned(16))) uint16_t As[BM][LDK];
  __shared__ __attribute__((aligned(16))) uint16_t Bs[BN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w >> 1, wn = w & 1;
  const int m0 = blockIdx.y * BM, n0 = blockIdx.x * BN;
  f32x16 acc;
#pragma unroll
  for (int i = 0; i < 16; ++i) acc[i] = 0.0f;
  for (int k0 = 0; k0 < K; k0 += BK) {
    // A tile
    if (A_KC) {
      const int m = tid >> 2, kc = (tid & 3) * 8;
      const int gm = m0 + m;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int gk = k0 + kc + i;
        As[m][kc + i] = (gm < M && gk < K) ? A[(size_t)gm * lda + gk] : (uint16_t)0;
      }
    } else {
      const int k = tid >> 3, mc = (tid & 7) * 8;
      const int gk = k0 + k;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int gm = m0 + mc + i;
        As[mc + i][k] = (gm < M && gk < K) ? A[(size_t)gk * lda + gm] : (uint16_t)0;
      }
    }
    // B tile
    if (B_KC) {
      const int n = tid >> 2, kc = (tid & 3) * 8;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int gk = k0 + kc + i;
        Bs[n][kc + i] = gk < K ? B[(size_t)(n0 + n) * ldb + gk] : (uint16_t)0;
      }
    } else {
      const int k = tid >> 3, nc = (tid & 7) * 8;
      const int gk = k0 + k;
#pragma unroll
      for (int i = 0; i < 8; ++i) Bs[nc + i][k] = gk < K ? B[(size_t)gk * ldb + n0 + nc + i] : (uint16_t)0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < BK; kk += 16) {
      const bf16x8 a = *reinterpret_cast<const bf16x8*>(&As[wm * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      const bf16x8 b = *reinterpret_cast<const bf16x8*>(&Bs[wn * 32 + (lane & 31)][kk + 8 * (lane >> 5)]);
      acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
    }
    __syncthreads();
  }
  const int col = n0 + wn * 32 + (lane & 31);
  const float bv = bias ? bias[col] : 0.0f;
#pragma unroll
  for (int i = 0; i < 16; ++i) {
    const int row = m0 + wm * 32 + 8 * (i >> 2) + 4 * (lane >> 5) + (i & 3);
    if (row < M) C[(size_t)row * ldc + col] = acc[i] + bv;
  }
}

// out[v*ldo + f] (+)= sum_{e in row_ptr[v]..row_ptr[v+1]} w[e] * in[col[e]*ldi + f]
// for f < H.  One wave per node, 4 features per lane; edges in CSR order.
__global__ __launch_bounds__(256) void k_aggregate(int nnodes, const int32_t* __restrict__ row_ptr,
                                                   const int32_t* __restrict__ col, const float* __restrict__ w,
                                                   const float* __restrict__ in, int ldi, float* __restrict__ out,
                                                   int ldo, int accumulate) {
  const int v = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (v >= nnodes) return;
  const int f = (threadIdx.x & 63) * 4;
  float4 s = make_float4(0.f, 0.f, 0.f, 0.f);
  for (int e = row_ptr[v]; e < row_ptr[v + 1]; ++e) {
    const float we = w[e];
    const float4 x = *reinterpret_cast<const float4*>(in + (size_t)col[e] * ldi + f);
    s.x += we * x.x;
    s.y += we * x.y;
    s.z += we * x.z;
    s.w += we * x.w;
  }
  float4* o = reinterpret_cast<float4*>(out + (size_t)v * ldo + f);
  if (accumulate) {
    const float4 a = *o;
    s.x += a.x;
    s.y += a.y;
    s.z += a.z;
    s.w += a.w;
  }
  *o = s;
}

// X[v] = [bf16(h[v]) | bf16(m[v])]  (rows of 2H).
__global__ void k_pack_x(int n, const float* __restrict__ h, const float* __restrict__ m, uint16_t* __restrict__ X) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (size_t)n * 2 * H) return;
  const size_t v = i / (2 * H), f = i % (2 * H);
  X[i] = f2bf(f < H ? h[v * H + f] : m[v * H + f - H]);
}

// h_out = (h_in ? h_in : 0) + max(z, 0)
__global__ void k_relu_add(size_t n, const float* __restrict__ z, const float* __restrict__ h_in,
                           float* __restrict__ h_out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  h_out[i] = (h_in ? h_in[i] : 0.0f) + fmaxf(z[i], 0.0f);
}

// g[gi] = mean of h over the graph's nodes; Y[gi] = [bf16(g) | bf16(gfeat)].
__global__ __launch_bounds__(256) void k_pool_head_in(const int32_t* __restrict__ node_off,
                                                      const float* __restrict__ h, const float* __restrict__ gfeat,
                                                      uint16_t* __restrict__ Y) {
  const int gi = blockIdx.x, f = threadIdx.x;
  const int a = node_off[gi], b = node_off[gi + 1];
  float s = 0.0f;
  for (int v = a; v < b; ++v) s += h[(size_t)v * H + f];
  Y[(size_t)gi * HIN + f] = f2bf(s / (float)(b - a));
  if (f < GF) Y[(size_t)gi * HIN + H + f] = f2bf(gfeat[gi * GF + f]);
}

// Ur = bf16(max(z, 0)) ; R = max(z, 0) kept in FP32 only through the mask.
__global__ void k_relu_bf16(size_t n, const float* __restrict__ z, uint16_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(fmaxf(z[i], 0.0f));
}

__global__ void k_to_bf16(size_t n, const float* __restrict__ x, uint16_t* __restrict__ out) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = f2bf(x[i]);
}

// Masked MSE: dO = 2 (O - T) mask / (2 sum p), per-graph loss partials.
__global__ void k_loss(int G, const float* __restrict__ O, const float* __restrict__ T, const int32_t* __restrict__ p,
                       float inv_denom, float* __restrict__ dO, uint16_t* __restrict__ dOr, float* __restrict__ part) {
  const int gi = blockIdx.x, c = threadIdx.x;  // 64 threads
  const int pg = p[gi];
  const bool used = (c < PMAX && c < pg) || (c >= PMAX && c < PMAX + pg);
  const float d = used ? O[gi * NOUT + c] - T[gi * 2 * PMAX + c] : 0.0f;
  const float g = 2.0f * d * inv_denom;
  dO[gi * NOUT + c] = g;
  dOr[gi * NOUT + c] = f2bf(g);
  __shared__ float s[64];
  s[c] = d * d;
  __syncthreads();
  if (c == 0) {
    float t = 0.0f;
    for (int k = 0; k < 64; ++k) t += s[k];
    part[gi] = t * inv_denom;
  }
}

// d = up * (z > 0); d_bf = bf16(d)
__global__ void k_relu_grad(size_t n, const float* __restrict__ up, int ldu, int width, const float* __restrict__ z,
                            float* __restrict__ d, uint16_t* __restrict__ dbf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t r = i / width, c = i % width;
  const float v = z[i] > 0.0f ? up[r * ldu + c] : 0.0f;
  d[i] = v;
  dbf[i] = f2bf(v);
}

// dh[v] = dY[graph(v)][0..H) / n_graph
__global__ void k_unpool(int G, const int32_t* __restrict__ node_off, const float* __restrict__ dY, int ldy,
                         float* __restrict__ dh) {
  const int gi = blockIdx.x, f = threadIdx.x;
  const int a = node_off[gi], b = node_off[gi + 1];
  const float v = dY[(size_t)gi * ldy + f] / (float)(b - a);
  for (int u = a; u < b; ++u) dh[(size_t)u * H + f] = v;
}

// dh[v][f] += dX[v][f] (f < H)
__global__ void k_add_left(size_t n, const float* __restrict__ dX, float* __restrict__ dh) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const size_t v = i / H, f = i % H;
  dh[i] += dX[v * 2 * H + f];
}

// out[c] = sum over rows of x[r*ld + c] (fixed order: deterministic).
__global__ void k_colsum(int rows, int cols, const float* __restrict__ x, int ld, float* __restrict__ out) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= cols) return;
  float s = 0.0f;
  for (int r = 0; r < rows; ++r) s += x[(size_t)r * ld + c];
  out[c] = s;
}

// Adam on the flat FP32 parameters, then their bfloat16 copies.
__global__ void k_adam(size_t n, float* __restrict__ prm, const float* __restrict__ g, float* __restrict__ m1,
                       float* __restrict__ m2, float lr, float b1, float b2, float eps, float c1, float c2,
                       uint16_t* __restrict__ pbf) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float a = b1 * m1[i] + (1.0f - b1) * gi;
  const float b = b2 * m2[i] + (1.0f - b2) * gi * gi;
  m1[i] = a;
  m2[i] = b;
  const float p = prm[i] - lr * (a / c1) / (sqrtf(b / c2) + eps);
  prm[i] = p;
  pbf[i] = f2bf(p);
}

inline unsigned blocks(size_t n, int t = 256) { return (unsigned)((n + t - 1) / t); }

}  // namespace

void gemm(hipStream_t st, bool a_kc, bool b_kc, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
          int ldb, float* C, int ldc, const float* bias) {
  const dim3 g(N / BN, (M + BM - 1) / BM), t(256);
  if (a_kc && !b_kc) hipLaunchKernelGGL((k_gemm_bf16<true, false>), g, t, 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias);
  else if (a_kc && b_kc) hipLaunchKernelGGL((k_gemm_bf16<true, true>), g, t, 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias);
  else if (!a_kc && !b_kc)
    hipLaunchKernelGGL((k_gemm_bf16<false, false>), g, t, 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias);
  else hipLaunchKernelGGL((k_gemm_bf16<false, true>), g, t, 0, st, M, N, K, A, lda, B, ldb, C, ldc, bias);
}

void aggregate(hipStream_t st, int nnodes, const int32_t* row_ptr, const int32_t* col, const float* w,
               const float* in, int ldi, float* out, int ldo, bool accumulate) {
  hipLaunchKernelGGL(k_aggregate, dim3((nnodes + 3) / 4), dim3(256), 0, st, nnodes, row_ptr, col, w, in, ldi, out, ldo,
                     accumulate ? 1 : 0);
}

void pack_x(hipStream_t st, int n, const float* h, const float* m, uint16_t* X) {
  hipLaunchKernelGGL(k_pack_x, dim3(blocks((size_t)n * 2 * H)), dim3(256), 0, st, n, h, m, X);
}

void relu_add(hipStream_t st, size_t n, const float* z, const float* h_in, float* h_out) {
  hipLaunchKernelGGL(k_relu_add, dim3(blocks(n)), dim3(256), 0, st, n, z, h_in, h_out);
}

void pool_head_in(hipStream_t st, int G, const int32_t* node_off, const float* h, const float* gfeat, uint16_t* Y) {
  hipLaunchKernelGGL(k_pool_head_in, dim3(G), dim3(H), 0, st, node_off, h, gfeat, Y);
}

void relu_bf16(hipStream_t st, size_t n, const float* z, uint16_t* out) {
  hipLaunchKernelGGL(k_relu_bf16, dim3(blocks(n)), dim3(256), 0, st, n, z, out);
}

void to_bf16(hipStream_t st, size_t n, const float* x, uint16_t* out) {
  hipLaunchKernelGGL(k_to_bf16, dim3(blocks(n)), dim3(256), 0, st, n, x, out);
}

void loss(hipStream_t st, int G, const float* O, const float* T, const int32_t* p, float inv_denom, float* dO,
          uint16_t* dOr, float* part) {
  hipLaunchKernelGGL(k_loss, dim3(G), dim3(64), 0, st, G, O, T, p, inv_denom, dO, dOr, part);
}

void relu_grad(hipStream_t st, int rows, int width, const float* up, int ldu, const float* z, float* d, uint16_t* dbf) {
  const size_t n = (size_t)rows * width;
  hipLaunchKernelGGL(k_relu_grad, dim3(blocks(n)), dim3(256), 0, st, n, up, ldu, width, z, d, dbf);
}

void unpool(hipStream_t st, int G, const int32_t* node_off, const float* dY, int ldy, float* dh) {
  hipLaunchKernelGGL(k_unpool, dim3(G), dim3(H), 0, st, G, node_off, dY, ldy, dh);
}

void add_left(hipStream_t st, int nnodes, const float* dX, float* dh) {
  const size_t n = (size_t)nnodes * H;
  hipLaunchKernelGGL(k_add_left, dim3(blocks(n)), dim3(256), 0, st, n, dX, dh);
}

void colsum(hipStream_t st, int rows, int cols, const float* x, int ld, float* out) {
  hipLaunchKernelGGL(k_colsum, dim3(blocks(cols, 64)), dim3(64), 0, st, rows, cols, x, ld, out);
}

void adam(hipStream_t st, size_t n, float* prm, const float* g, float* m1, float* m2, float lr, float b1, float b2,
          float eps, float c1, float c2, uint16_t* pbf) {
  hipLaunchKernelGGL(k_adam, dim3(blocks(n)), dim3(256), 0, st, n, prm, g, m1, m2, lr, b1, b2, eps, c1, c2, pbf);
}

}  // namespace mpnn
}  // namespace rq
