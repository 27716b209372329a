// Grover search (crates/ruqu-algorithms/src/grover.rs) on the device, and the
// amplitude sign flips it performs through QuantumState::amplitudes_mut.
//
// The reference's sequence is kept gate for gate so that amplitudes are
// bit-identical: H on every qubit, then per iteration the oracle (negate each
// listed target, grover.rs:130-139), H on every qubit, negate every amplitude
// but index 0 (:146-155), H on every qubit; success probability = sum of the
// targets' probabilities in list order (:161-162); measured state = qubits
// 0..n-1 measured in turn (:187-200), with the draws as inputs.
#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "fusion.h"
#include "internal.h"

namespace rq {
namespace {

// Distinct indices whose listed count is odd (an index listed twice is
// negated twice by the reference's loop, i.e. left unchanged).
std::vector<uint64_t> odd_indices(const uint64_t* idx, size_t count) {
  std::vector<uint64_t> v(idx, idx + count), out;
  std::sort(v.begin(), v.end());
  for (size_t i = 0; i < v.size();) {
    size_t j = i;
    while (j < v.size() && v[j] == v[i]) ++j;
    if ((j - i) & 1) out.push_back(v[i]);
    i = j;
  }
  return out;
}

int hadamard_all(State& s) {
  for (uint32_t q = 0; q < s.nq; ++q) {
    rq_gate g{};
    g.kind = RQ_GATE_H;
    g.q0 = q;
    int has = 0;
    rq_measurement m;
    if (int rc = apply_gate_impl(s, g, 0.0, &m, &has)) return rc;
  }
  return RQ_OK;
}

double prob_at(const State& s, uint64_t i, int* rc) {  // probabilities()[i] (state.rs:123-126)
  if (s.prec == RQ_C64) {
    double a[2];
    hipError_t e = hipMemcpy(a, (const double*)s.amps + 2 * i, sizeof(a), hipMemcpyDeviceToHost);
    if (e != hipSuccess) *rc = hip_fail(e, "grover");
    return a[0] * a[0] + a[1] * a[1];
  }
  float a[2];
  hipError_t e = hipMemcpy(a, (const float*)s.amps + 2 * i, sizeof(a), hipMemcpyDeviceToHost);
  if (e != hipSuccess) *rc = hip_fail(e, "grover");
  return (double)(a[0] * a[0] + a[1] * a[1]);  // mixed_precision.rs:303-305
}

}  // namespace
}  // namespace rq

using namespace rq;

extern "C" {

uint32_t rq_grover_optimal_iterations(uint32_t nq, size_t ntargets) {  // grover.rs:69-74
  const double n = std::ldexp(1.0, (int)nq);
  const double theta = std::asin(std::sqrt((double)ntargets / n));
  const double k = std::fmax(std::round(M_PI / 4.0 / theta - 0.5), 1.0);  // f64::max ignores NaN, so does fmax
  if (!(k < 4294967295.0)) return 4294967295u;  // Rust's saturating `as u32`
  return (uint32_t)k;
}

int rq_negate_amplitudes(rq_state* h, const uint64_t* indices, size_t count) {
  State* s = (State*)h;
  if (!s || (count && !indices)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  for (size_t k = 0; k < count; ++k)
    if (indices[k] >= s->n)
      return fail(RQ_ERR_INVALID_STATE, "amplitude index " + std::to_string(indices[k]) + " out of range for " +
                                            std::to_string(s->n) + " amplitudes");
  std::vector<uint64_t> odd = odd_indices(indices, count);
  if (odd.empty()) return RQ_OK;
  uint64_t* d = nullptr;
  RQ_HIP(hipMalloc((void**)&d, odd.size() * sizeof(uint64_t)));
  hipError_t e = hipMemcpyAsync(d, odd.data(), odd.size() * sizeof(uint64_t), hipMemcpyHostToDevice, s->stream);
  if (e == hipSuccess) {
    launch_negate_listed(*s, d, odd.size());
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  return e == hipSuccess ? RQ_OK : hip_fail(e, "negate_amplitudes");
}

int rq_negate_amplitudes_from(rq_state* h, uint64_t start) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  launch_negate_from(*s, start);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

int rq_grover_run(uint32_t nq, const uint64_t* targets, size_t ntargets, int64_t iterations, int precision,
                  int with_seed, const double* draws, rq_grover_result* out, rq_state** state_out) {
  if (!out || (ntargets && !targets) || (nq && !draws)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (state_out) *state_out = nullptr;
  const uint64_t dim = nq < 64 ? (uint64_t)1 << nq : 0;
  for (size_t k = 0; k < ntargets; ++k)  // grover.rs:104-111 (an assert in the reference)
    if (targets[k] >= dim)
      return fail(RQ_ERR_ARGUMENT, "target state index " + std::to_string(targets[k]) + " out of range for " +
                                       std::to_string(nq) + " qubits (max " + std::to_string(dim - 1) + ")");
  if (ntargets == 0 && iterations < 0)
    return fail(RQ_ERR_ARGUMENT, "no target states: the optimal iteration count is unbounded");
  const uint32_t iters = iterations < 0 ? rq_grover_optimal_iterations(nq, ntargets) : (uint32_t)iterations;
  rq_state* hs = nullptr;
  if (int rc = rq_state_create(nq, precision, with_seed, &hs)) return rc;
  State& s = *(State*)hs;
  std::vector<uint64_t> odd = odd_indices(targets, ntargets);
  uint64_t* d = nullptr;
  int rc = RQ_OK;
  if (!odd.empty()) {
    hipError_t e = hipMalloc((void**)&d, odd.size() * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMemcpy(d, odd.data(), odd.size() * sizeof(uint64_t), hipMemcpyHostToDevice);
    if (e != hipSuccess) rc = hip_fail(e, "grover");
  }
  if (!rc) rc = hadamard_all(s);
  for (uint32_t it = 0; it < iters && !rc; ++it) {
    if (d) launch_negate_listed(s, d, odd.size());
    rc = hadamard_all(s);
    if (!rc) launch_negate_from(s, 1);
    if (!rc) rc = hadamard_all(s);
  }
  if (!rc) {
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) rc = hip_fail(e, "grover");
  }
  double psucc = 0.0;
  for (size_t k = 0; k < ntargets && !rc; ++k) psucc += prob_at(s, targets[k], &rc);
  uint64_t measured = 0;
  for (uint32_t q = 0; q < nq && !rc; ++q) {
    rq_measurement m;
    rc = measure_state(s, q, draws[q], &m);
    if (!rc && m.result) measured |= (uint64_t)1 << q;
  }
  if (d) (void)hipFree(d);
  if (rc) {
    rq_state_destroy(hs);
    return rc;
  }
  out->measured_state = measured;
  out->target_found = std::find(targets, targets + ntargets, measured) != targets + ntargets ? 1 : 0;
  out->success_probability = psucc;
  out->num_iterations = iters;
  if (state_out) *state_out = hs;
  else rq_state_destroy(hs);
  return RQ_OK;
}

}  // extern "C"
