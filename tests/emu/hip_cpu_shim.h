// Host emulation of the HIP device features used by the batched kernels (test
// infrastructure): the 256 work-items of a block are cooperative contexts run
// by one host thread and switched at barriers (block barrier for
// __syncthreads, a 64-item barrier for readfirstlane), shuffles go through a
// per-block exchange array, and blocks run one after another.  Lets the
// kernel source run on the CPU without a GPU.
#pragma once
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <cmath>
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
extern Dim3 threadIdx, blockIdx, blockDim, gridDim;

namespace emu {
struct Barrier {
  int count = 0, target = 0;
  unsigned gen = 0;
};
void wait(Barrier& b);  // yields to the other work-items until all arrived
extern Barrier block_bar, wave_bar[4];
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

inline void __syncthreads() { emu::wait(emu::block_bar); }

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
  emu::wait(emu::wave_bar[w]);
  const int r = emu::wx[w];
  emu::wait(emu::wave_bar[w]);
  return r;
}
// The hardware sine / cosine take the angle in turns.
inline float __builtin_amdgcn_sinf(float r) { return (float)std::sin(6.283185307179586476925 * (double)r); }
inline float __builtin_amdgcn_cosf(float r) { return (float)std::cos(6.283185307179586476925 * (double)r); }
inline int max(int a, int b) { return a > b ? a : b; }
inline int min(int a, int b) { return a < b ? a : b; }
inline int __popc(uint32_t x) { return __builtin_popcount(x); }
inline int __popcll(uint64_t x) { return __builtin_popcountll(x); }
inline unsigned atomicAdd(unsigned* p, unsigned v) { return __atomic_fetch_add(p, v, __ATOMIC_SEQ_CST); }
inline unsigned long long atomicMax(unsigned long long* p, unsigned long long v) {
  const unsigned long long old = *p;
  if (v > old) *p = v;
  return old;
}
inline long long __double_as_longlong(double x) {
  long long r;
  memcpy(&r, &x, sizeof r);
  return r;
}
