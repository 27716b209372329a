// Internal declarations shared by the host runtime and the HIP kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/ruqu_gpu.h"

namespace rq {

// Thread-local last error (rq_last_error) and status helpers.
int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);

#define RQ_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return ::rq::hip_fail(_e, #call); \
  } while (0)

// One grid dimension may cover at most 2^32 - 1 work-items, which a
// one-thread-per-amplitude launch exceeds from 32 qubits on.  Element-wise
// launches therefore put at most 2^16 workgroups in a row and the rows in
// gridDim.y; kernels take their workgroup index from flat_block() and keep
// their bounds checks (the last row may run past the end).
constexpr uint64_t GRID_ROW = 1ull << 16;
inline dim3 grid_blocks(uint64_t blocks) {
  if (blocks <= GRID_ROW) return dim3((unsigned)(blocks ? blocks : 1));
  return dim3((unsigned)GRID_ROW, (unsigned)((blocks + GRID_ROW - 1) / GRID_ROW));
}
__device__ __forceinline__ uint64_t flat_block() { return (uint64_t)blockIdx.y * gridDim.x + blockIdx.x; }

struct FusedProgram;  // the fused program cached in State::prog (kernels_fused.hip)

// Device-resident state vector.
struct State {
  uint32_t nq = 0;
  int prec = RQ_C64;
  size_t n = 0;              // 2^nq amplitudes
  void* amps = nullptr;      // interleaved (re, im), 16 B or 8 B per amplitude
  hipStream_t stream = nullptr;
  double* red = nullptr;     // reduction partials (device)
  size_t red_cap = 0;        // doubles
  double* host_red = nullptr;  // pinned host mirror for small results
  std::vector<rq_measurement> record;
  uint64_t gate_count = 0;   // for precision_error_bound
  void* samp = nullptr;      // shot sampler's tree, draws and bits, kept when at most SAMPLER_KEEP_BYTES
  size_t samp_cap = 0;
  void* prog = nullptr;      // device copy of the fused / dense program
  size_t prog_cap = 0;
  FusedProgram* fused = nullptr;  // what `prog` holds when it is a reusable fused program
  size_t amp_bytes() const { return prec == RQ_C64 ? 16 : 8; }
  size_t bytes() const { return n * amp_bytes(); }
};

// One gate with the exact kernels (state.rs:158-204 semantics).
int apply_gate_impl(State& s, const rq_gate& g, double draw, rq_measurement* out, int* has_out);
// Algorithmic bytes of the exact kernel for a gate kind.
double gate_bytes(const State& s, int kind);

// The fused program cached in s.prog: invalidate before writing s.prog
// otherwise; release when the state is destroyed.
void invalidate_fused_program(State& s);
void release_fused_program(State& s);

// Ensure at least `doubles` of reduction scratch on the state.
int ensure_red(State& s, size_t doubles);

// Shape knobs of the streaming (one pass per gate) kernels.
struct StreamTune {
  int unroll = 1;          // amplitude groups per thread: 1, 2, 4 or 8
  bool nontemporal = true; // measured best on MI355X (30-qubit c32 sweep)
  int xcd_chunk = -1;      // two-stream 1q kernels' block order: -1 per-distance table, 0 linear, else chunk
};
StreamTune& stream_tune();

// ---- kernel launchers (kernels_gates.hip / kernels_reduce.hip) ----------
// m: row-major 2x2 / 4x4 complex-f64 matrices; converted to the state's
// precision on the host exactly as mixed_precision.rs:556-591 does.
void launch_init_zero(State& s);
void launch_apply_1q(State& s, uint32_t q, const rq_c64* m);
void launch_apply_2q(State& s, uint32_t q1, uint32_t q2, const rq_c64* m);
void launch_cnot(State& s, uint32_t c, uint32_t t);
void launch_cz(State& s, uint32_t q1, uint32_t q2);
void launch_swap(State& s, uint32_t q1, uint32_t q2);
void launch_diag_2q(State& s, uint32_t q1, uint32_t q2, const rq_c64* d4);
void launch_collapse(State& s, uint32_t q, int result, double nf64);
void launch_scale(State& s, double f);
void launch_negate_from(State& s, uint64_t start);                       // a[i] = -a[i], i >= start
void launch_negate_listed(State& s, const uint64_t* d_idx, uint64_t count);  // distinct device indices
void launch_copy(State& dst, const State& src);
// dst and src of the same size and opposite precisions: narrow (to c32) or widen (to c64).
void launch_convert(State& dst, const State& src);
void launch_plain_copy(void* dst, const void* src, size_t bytes, hipStream_t st);

// Reductions: results land in s.red (device); the caller reads them back.
// Sum over amplitudes with (i & mask) == want of |a_i|^2 (FP64 accumulate).
int reduce_masked_normsq(State& s, uint64_t mask, uint64_t want, double* out);
// P(q = 0) of a complex-f32 state summed sequentially in f32 in index order,
// bit for bit as mixed_precision.rs:480-486 (one workgroup).
int reduce_p0_seq_f32(State& s, uint32_t q, float* out);
// Per-term <P_k> for canonicalised Pauli strings (x, zy parity mask, nY mod 4).
struct PauliCanon {
  uint64_t x;
  uint64_t zy;
  uint32_t ny;
};
int reduce_pauli(State& s, const PauliCanon* terms, size_t nterms, double* out);
// Allocation-free: one batch of 1..32 terms, partials in a fixed device
// scratch array; adds coeffs[k] * <P_k> to *acc in term order.
int reduce_hamiltonian_noalloc(State& s, const double* coeffs, const PauliCanon* terms, size_t nterms, double* acc);
int reduce_z_all(State& s, double* out);
int reduce_inner(const State& a, const State& b, double* re, double* im);
int launch_probabilities(State& s, double* d_out);
int sample_shots(State& s, const double* draws, size_t shots, uint8_t* bits, uint64_t* near_ties);

}  // namespace rq
