// CPU run of the batched kernel source (tests only): the kernel file is
// included after the shim; launches run block by block with 256 threads.
#include "hip_cpu_shim.h"

thread_local Dim3 threadIdx, blockIdx;
Dim3 blockDim, gridDim;
namespace emu {
std::barrier<>* bar = nullptr;
std::barrier<>* wbar[4] = {};
int wx[4] = {};
std::vector<double> xchg;
unsigned char* lds = nullptr;
}  // namespace emu

#include "batch_kernel_src.inc"

extern "C" {

// Runs k_batch_eval with `grid` blocks (sequentially, 256 threads each).
int emu_batch_eval(const rq::BInst* insts, const rq::BItem* items, int nitems, const int32_t* done,
                   const rq::BEdge* edges, const rq::BTerm* terms, const double* params, const float* ctab,
                   float2* slots, uint64_t slot_amps, double* energy, uint32_t* amax, int grid) {
  const int B = rq::BATCH_THREADS;
  unsigned counter = 0;
  std::vector<unsigned char> lds((1 << rq::BATCH_LDS_QUBITS) * 8);
  emu::lds = lds.data();
  emu::xchg.assign(B, 0.0);
  blockDim.x = B;
  gridDim.x = grid;
  for (int b = 0; b < grid; ++b) {
    std::barrier<> bar(B);
    emu::bar = &bar;
    std::barrier<> w0(64), w1(64), w2(64), w3(64);
    emu::wbar[0] = &w0;
    emu::wbar[1] = &w1;
    emu::wbar[2] = &w2;
    emu::wbar[3] = &w3;
    std::vector<std::thread> th;
    for (int t = 0; t < B; ++t)
      th.emplace_back([&, t, b] {
        threadIdx.x = t;
        blockIdx.x = b;
        rq::k_batch_eval(insts, items, nitems, &counter, done, edges, terms, params, ctab, slots, slot_amps, energy,
                         amax);
      });
    for (auto& x : th) x.join();
  }
  return 0;
}

void emu_cut_table(float* ctab, int n, const rq::BEdge* e, int ne) {
  threadIdx.x = 0;
  blockIdx.x = 0;
  blockDim.x = 1;
  gridDim.x = 1;
  const uint64_t N = 1ull << n;
  for (uint64_t i = 0; i < N; ++i) ctab[i] = (float)rq::emu_cut_of(i, e, ne);
}
}
