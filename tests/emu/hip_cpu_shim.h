// Host-thread emulation of the HIP device features used by the batched
// kernel: one std::thread per work-item of a block, a real barrier for
// __syncthreads, per-wave barriers for readfirstlane, shuffles through a
// per-block exchange array.  Lets the
// kernel source run on the CPU for tests (no GPU involved).
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <barrier>
#include <cmath>
#include <thread>
#include <vector>

struct float2 {
  float x, y;
};
struct double2 {
  double x, y;
};
inline float2 make_float2(float x, float y) { return float2{x, y}; }

struct Dim3 {
  unsigned x = 1, y = 1, z = 1;
};
extern thread_local Dim3 threadIdx, blockIdx;
extern Dim3 blockDim, gridDim;

namespace emu {
extern std::barrier<>* bar;
extern std::barrier<>* wbar[4];  // one per 64-thread wave of the block
extern int wx[4];
extern std::vector<double> xchg;
extern unsigned char* lds;
}  // namespace emu

#define __global__
#define __device__
#define __host__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __restrict__

inline void __syncthreads() { emu::bar->arrive_and_wait(); }

inline double __shfl_down(double v, int off, int width) {
  unsigned t = threadIdx.x;
  emu::xchg[t] = v;
  __syncthreads();
  unsigned lane = t % width;
  double r = (lane + off < (unsigned)width) ? emu::xchg[t + off] : v;
  __syncthreads();
  return r;
}
// Lane 0's value, for every thread of the calling wave (the whole wave calls
// it, as in the kernel's uniform code).
inline int __builtin_amdgcn_readfirstlane(int v) {
  const unsigned w = threadIdx.x / 64;
  if (threadIdx.x % 64 == 0) emu::wx[w] = v;
  emu::wbar[w]->arrive_and_wait();
  const int r = emu::wx[w];
  emu::wbar[w]->arrive_and_wait();
  return r;
}
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
