// Dense 5-qubit blocks on the matrix cores (complex-f32 states).
//
// A block is a 32x32 complex unitary U on qubits q0 < ... < q4.  Every group
// of 32 amplitudes that differ only in those qubits is transformed y = U x.
// With the real embedding W = [[Re U, -Im U], [Im U, Re U]] (64 x 64) the
// groups form one real GEMM Y^T = W X^T (columns = groups, 64 rows = real and
// imaginary parts), run on v_mfma_f32_32x32x2_f32: a wave takes 32 groups
// (1024 amplitudes) at a time, 32 K-steps x 2 row halves = 64 matrix-core
// instructions, 256 flop per amplitude, FP32 accumulation.
//
// Operand layout of v_mfma_f32_32x32x2_f32: lane l holds A[l % 32][l / 32]
// and B[l / 32][l % 32]; its 16 results are column l % 32, rows
// 8*(i/4) + 4*(l/32) + i%4.  With A = W and B = X^T, lane l reads, for group
// l % 32, the 16 amplitudes whose block index c has c % 2 == l / 32 (their
// real parts feed K-steps 0..15, their imaginary parts steps 16..31), and
// writes 16 output amplitudes of the same group: both the loads and the
// stores of a wave's 32 lanes run over consecutive groups (contiguous
// whenever the lowest qubits are outside the block).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "internal.h"

namespace rq {
namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));

struct BlockQ {
  int q[5];
};

__device__ __forceinline__ uint64_t spread5(uint64_t g, const BlockQ& b) {
#pragma unroll
  for (int i = 0; i < 5; ++i) {
    const int q = b.q[i];
    g = ((g >> q) << (q + 1)) | (g & ((1ull << q) - 1ull));
  }
  return g;
}

__device__ __forceinline__ uint64_t offset5(int c, const BlockQ& b) {
  uint64_t o = 0;
#pragma unroll
  for (int i = 0; i < 5; ++i) o |= (uint64_t)((c >> i) & 1) << b.q[i];
  return o;
}

// wt[(h*32 + s)*64 + l] = W[32h + l%32][2s + l/32] (the A operand of lane l
// for K-step s and row half h), prepared on the host.  The block keeps it in
// LDS (16 KB, read once per MFMA step), which leaves the registers to two
// tiles of amplitudes: each wave loads its next tile while the matrix cores
// work on the current one.
__global__ __launch_bounds__(256) void k_dense5_f32(float2* __restrict__ a, uint64_t ntiles, BlockQ bq,
                                                    const float* __restrict__ wt) {
  __shared__ float wl[64 * 64];
  for (int i = threadIdx.x; i < 64 * 64; i += 256) wl[i] = wt[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, row = lane & 31, half = lane >> 5;
  const uint64_t wave = (uint64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const uint64_t nwaves = (uint64_t)gridDim.x * 4;
  // block offsets are linear in the block index: a wave-uniform part plus
  // the lane half's bit (bit 0 for the inputs, bit 2 for the outputs)
  const uint64_t hin = half ? (1ull << bq.q[0]) : 0ull, hout = half ? (1ull << bq.q[2]) : 0ull;
  float2 xn[16];
  if (wave < ntiles) {
    const uint64_t base = spread5(wave * 32 + row, bq) + hin;
#pragma unroll
    for (int j = 0; j < 16; ++j) xn[j] = a[base + offset5(2 * j, bq)];
  }
  for (uint64_t tile = wave; tile < ntiles; tile += nwaves) {
    const uint64_t base = spread5(tile * 32 + row, bq);
    float2 x[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) x[j] = xn[j];
    const uint64_t next = tile + nwaves;
    if (next < ntiles) {
      const uint64_t nb = spread5(next * 32 + row, bq) + hin;
#pragma unroll
      for (int j = 0; j < 16; ++j) xn[j] = a[nb + offset5(2 * j, bq)];
    }
    f32x16 acc0, acc1;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      acc0[i] = 0.0f;
      acc1[i] = 0.0f;
    }
    // W is re-read from LDS for every tile: a per-tile zero keeps the
    // compiler from hoisting the 64 values back into registers
    const float* __restrict__ wr = wl + lane + __builtin_amdgcn_readfirstlane((int)(tile >> 62));
#pragma unroll
    for (int s = 0; s < 32; ++s) {
      const float xv = s < 16 ? x[s].x : x[s - 16].y;
      acc0 = __builtin_amdgcn_mfma_f32_32x32x2f32(wr[s * 64], xv, acc0, 0, 0, 0);
      acc1 = __builtin_amdgcn_mfma_f32_32x32x2f32(wr[(32 + s) * 64], xv, acc1, 0, 0, 0);
    }
    // acc0[i] / acc1[i]: real / imaginary part of output amplitude
    // c = 8 (i / 4) + 4 half + i % 4 of this lane's group
#pragma unroll
    for (int i = 0; i < 16; ++i) a[base + hout + offset5(8 * (i >> 2) + (i & 3), bq)] = make_float2(acc0[i], acc1[i]);
  }
}

}  // namespace

// Applies one dense 5-qubit block (qubits ascending, wt as above, resident on
// the device) to a complex-f32 state of at least 10 qubits.
void launch_dense5_f32(State& s, const int* q, const float* d_wt) {
  BlockQ bq;
  for (int i = 0; i < 5; ++i) bq.q[i] = q[i];
  const uint64_t ntiles = (uint64_t)s.n >> 10;  // 32 groups of 32 amplitudes per wave-tile
  uint64_t blocks = (ntiles + 3) / 4;
  if (blocks > 2048) blocks = 2048;
  hipLaunchKernelGGL(k_dense5_f32, dim3((unsigned)blocks), dim3(256), 0, s.stream, (float2*)s.amps, ntiles, bq, d_wt);
}

}  // namespace rq
