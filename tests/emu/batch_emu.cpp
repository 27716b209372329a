// CPU run of the batched kernels' source (tests only): the kernel file is
// included after the shim; each launch runs its blocks one after another,
// with the 256 work-items of a block as cooperative contexts.
#include <ucontext.h>

#include <functional>

#include "hip_cpu_shim.h"

Dim3 threadIdx, blockIdx, blockDim, gridDim;
namespace emu {
Barrier block_bar, wave_bar[4];
int wx[4] = {};
std::vector<double> xchg;
unsigned char* lds = nullptr;

namespace {
ucontext_t sched_ctx;
ucontext_t* cur_ctx = nullptr;
bool cur_finished = false;
const std::function<void()>* cur_body = nullptr;
void trampoline() {
  (*cur_body)();
  cur_finished = true;  // uc_link returns to the scheduler
}
}  // namespace

void wait(Barrier& b) {
  const unsigned g = b.gen;
  if (++b.count == b.target) {
    b.count = 0;
    ++b.gen;
    return;
  }
  while (b.gen == g) swapcontext(cur_ctx, &sched_ctx);
}
}  // namespace emu

#include "batch_kernel_src.inc"

namespace {

std::vector<unsigned char> g_lds(((1 << rq::BATCH_MAX_TILE_QUBITS) + 256) * 8);
constexpr size_t STACK = 128 << 10;
std::vector<char> g_stacks((size_t)rq::BATCH_THREADS * STACK);

// Runs every block of the launch in order; the block's 256 work-items are
// contexts switched at barriers, so block-shared (static) data and the LDS
// belong to one block at a time.
void run_grid(int grid, const std::function<void()>& body) {
  const int B = rq::BATCH_THREADS;
  emu::lds = g_lds.data();
  emu::xchg.assign(B, 0.0);
  blockDim.x = B;
  gridDim.x = grid;
  gridDim.y = 1;
  blockIdx.y = 0;  // one-row grids
  emu::cur_body = &body;
  std::vector<ucontext_t> ctx(B);
  std::vector<char> fin(B);
  for (int b = 0; b < grid; ++b) {
    blockIdx.x = b;
    emu::block_bar = emu::Barrier{0, B, 0};
    for (auto& w : emu::wave_bar) w = emu::Barrier{0, 64, 0};
    for (int t = 0; t < B; ++t) {
      getcontext(&ctx[t]);
      ctx[t].uc_stack.ss_sp = g_stacks.data() + (size_t)t * STACK;
      ctx[t].uc_stack.ss_size = STACK;
      ctx[t].uc_link = &emu::sched_ctx;
      makecontext(&ctx[t], emu::trampoline, 0);
      fin[t] = 0;
    }
    int live = B;
    while (live) {
      for (int t = 0; t < B; ++t) {
        if (fin[t]) continue;
        threadIdx.x = t;
        emu::cur_ctx = &ctx[t];
        emu::cur_finished = false;
        swapcontext(&emu::sched_ctx, &ctx[t]);
        if (emu::cur_finished) {
          fin[t] = 1;
          --live;
        }
      }
    }
  }
}

}  // namespace

extern "C" {

int emu_batch_eval(const rq::BInst* insts, const rq::BItem* items, int nitems, const int32_t* done,
                   const rq::BEdge* edges, const rq::BTerm* terms, const double* params, const float* ctab,
                   double* energy, uint32_t* amax, int grid) {
  unsigned counter = 0;
  run_grid(grid, [&] {
    rq::k_batch_eval(insts, items, nitems, &counter, done, edges, terms, params, ctab, energy, amax);
  });
  return 0;
}

int emu_wide_pass(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wmap, const int32_t* wtile0,
                  const int64_t* wstate, const int32_t* done, const double* params, const float* ctab, float2* states,
                  int tq, int pass, rq::BPart* parts, const int32_t* ptile, const int32_t* first, const int32_t* last,
                  const int64_t* src) {
  const rq::WidePasses wp{ptile, first, last, src};
  run_grid(grid, [&] {
    switch (tq) {
      case 9: rq::k_wide_pass<9>(insts, wi, wmap, wtile0, wstate, wp, done, params, ctab, states, pass, parts); break;
      case 10: rq::k_wide_pass<10>(insts, wi, wmap, wtile0, wstate, wp, done, params, ctab, states, pass, parts); break;
      case 11: rq::k_wide_pass<11>(insts, wi, wmap, wtile0, wstate, wp, done, params, ctab, states, pass, parts); break;
      case 12: rq::k_wide_pass<12>(insts, wi, wmap, wtile0, wstate, wp, done, params, ctab, states, pass, parts); break;
      default: rq::k_wide_pass<13>(insts, wi, wmap, wtile0, wstate, wp, done, params, ctab, states, pass, parts); break;
    }
  });
  return 0;
}

int emu_wide_energy(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wmap,
                    const int32_t* wtile0, const int64_t* wstate, const int32_t* vmap, const int32_t* done,
                    const rq::BTerm* terms, const float2* states, rq::BPart* parts) {
  run_grid(grid, [&] { rq::k_wide_energy(insts, wi, wmap, wtile0, wstate, vmap, done, terms, states, parts); });
  return 0;
}

int emu_wide_fold(int grid, const rq::BInst* insts, const rq::BItem* wi, const int32_t* wtile0, const int32_t* done,
                  const rq::BPart* parts, double* energy, uint32_t* amax) {
  run_grid(grid, [&] { rq::k_wide_fold(insts, wi, wtile0, done, parts, energy, amax); });
  return 0;
}

int emu_cut_tables(float* ctab, const rq::BInst* insts, const rq::BEdge* edges, const rq::BCutChunk* chunks,
                   int nchunks) {
  run_grid(nchunks, [&] { rq::k_cut_tables(ctab, insts, edges, chunks, (uint32_t)nchunks); });
  return 0;
}
}
