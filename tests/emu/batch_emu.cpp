// CPU run of the batched kernels' source (tests only): the kernel file is
// included after the shim; each launch runs its blocks one after another,
// 256 host threads per block.
#include <functional>

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

namespace {

std::vector<unsigned char> g_lds((1 << rq::BATCH_MAX_TILE_QUBITS) * 8);

void run_grid(int grid, const std::function<void()>& body) {
  const int B = rq::BATCH_THREADS;
  emu::lds = g_lds.data();
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
        body();
      });
    for (auto& x : th) x.join();
  }
}

}  // namespace

extern "C" {

int emu_batch_eval(const rq::BInst* insts, const rq::BItem* items, int nitems, const int32_t* done,
                   const rq::BEdge* edges, const rq::BTerm* terms, const double* params, double* energy,
                   uint32_t* amax, int grid) {
  unsigned counter = 0;
  run_grid(grid, [&] { rq::k_batch_eval(insts, items, nitems, &counter, done, edges, terms, params, energy, amax); });
  return 0;
}

int emu_wide_a(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wmap, const int32_t* wtile0,
               const int64_t* wstate, const int32_t* done, const double* params, const float* ctab, float2* states,
               int tq, int layer) {
  run_grid(grid, [&] { rq::k_wide_a(insts, wi, wmap, wtile0, wstate, done, params, ctab, states, tq, layer); });
  return 0;
}

int emu_wide_b(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wmap, const int32_t* wtile0,
               const int64_t* wstate, const int32_t* done, const double* params, float2* states, int tq, int layer) {
  run_grid(grid, [&] { rq::k_wide_b(insts, wi, wmap, wtile0, wstate, done, params, states, tq, layer); });
  return 0;
}

int emu_wide_energy(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wmap,
                    const int32_t* wtile0, const int64_t* wstate, const int32_t* done, const rq::BTerm* terms,
                    const float* ctab, const float2* states, int tq, rq::BPart* parts) {
  run_grid(grid, [&] { rq::k_wide_energy(insts, wi, wmap, wtile0, wstate, done, terms, ctab, states, tq, parts); });
  return 0;
}

int emu_wide_fold(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wtile0, const int32_t* done,
                  const rq::BPart* parts, int tq, double* energy, uint32_t* amax) {
  run_grid(grid, [&] { rq::k_wide_fold(insts, wi, wtile0, done, parts, tq, energy, amax); });
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
