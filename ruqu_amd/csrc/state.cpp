// Host runtime behind the C ABI: QuantumState (state.rs), QuantumStateF32
// (mixed_precision.rs) and Simulator (simulator.rs) on device memory.
//
// Validation and error texts follow error.rs:8-23 and the checks in
// state.rs:31-96,158-204,441-449; every failure returns a status and sets the
// thread-local message, nothing aborts.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fusion.h"
#include "gates.h"
#include "internal.h"

namespace rq {

namespace {
thread_local std::string g_last_error;

int g_max_q64 = -1, g_max_q32 = -1;

int device_ready() {
  static int ready = -1;
  if (ready >= 0) return ready ? RQ_OK : fail(RQ_ERR_DEVICE, "no usable HIP device");
  int count = 0;
  hipError_t e = hipGetDeviceCount(&count);
  if (e != hipSuccess || count == 0) {
    ready = 0;
    return fail(RQ_ERR_DEVICE, "no usable HIP device");
  }
  ready = 1;
  return RQ_OK;
}

void compute_limits() {
  if (g_max_q64 >= 0) return;
  g_max_q64 = 34;
  g_max_q32 = 35;
  size_t free_b = 0, total_b = 0;
  int count = 0;
  if (hipGetDeviceCount(&count) == hipSuccess && count > 0 && hipMemGetInfo(&free_b, &total_b) == hipSuccess &&
      total_b > 0) {
    int q64 = 0, q32 = 0;
    while (q64 < 40 && ((size_t)16 << (q64 + 1)) <= total_b) ++q64;
    while (q32 < 40 && ((size_t)8 << (q32 + 1)) <= total_b) ++q32;
    g_max_q64 = q64;
    g_max_q32 = q32;
  }
}

int check_qubit(const State& s, uint32_t q) {  // state.rs:441-449
  if (q >= s.nq)
    return fail(RQ_ERR_INVALID_QUBIT,
                "invalid qubit index " + std::to_string(q) + " for " + std::to_string(s.nq) + "-qubit system");
  return RQ_OK;
}

int limit_error(uint32_t nq, int maxq) {  // error.rs:10-11
  return fail(RQ_ERR_QUBIT_LIMIT,
              "qubit limit exceeded: requested " + std::to_string(nq) + ", maximum " + std::to_string(maxq));
}

int alloc_state(uint32_t nq, int prec, State** out) {
  int rc = device_ready();
  if (rc) return rc;
  State* s = new State();
  s->nq = nq;
  s->prec = prec;
  s->n = (size_t)1 << nq;
  hipError_t e = hipMalloc(&s->amps, s->bytes());
  if (e != hipSuccess) {
    delete s;
    (void)hipGetLastError();
    return fail(RQ_ERR_ALLOC, "memory allocation failed: need " + std::to_string(((size_t)1 << nq) *
                                                                                 (prec == RQ_C64 ? 16 : 8)) +
                                  " bytes");
  }
  e = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipHostMalloc((void**)&s->host_red, 64 * sizeof(double), hipHostMallocDefault);
  if (e != hipSuccess) {
    if (s->stream) (void)hipStreamDestroy(s->stream);
    (void)hipFree(s->amps);
    delete s;
    return hip_fail(e, "state setup");
  }
  *out = s;
  return RQ_OK;
}

// Measurement with a supplied draw: state.rs:288-326 (f64: p0 summed in FP64)
// and mixed_precision.rs:473-515 (f32: p0 summed sequentially in f32 in index
// order, bit for bit, compared with the draw as f64, 1/sqrt in f32).
int measure_impl(State& s, uint32_t q, double draw, rq_measurement* out) {
  int rc = check_qubit(s, q);
  if (rc) return rc;
  int result;
  double prob, nf;
  if (s.prec == RQ_C64) {
    double p0d = 0.0;
    rc = reduce_masked_normsq(s, 1ull << q, 0, &p0d);
    if (rc) return rc;
    result = draw >= p0d;
    prob = result ? 1.0 - p0d : p0d;
    nf = prob > 0.0 ? 1.0 / std::sqrt(prob) : 0.0;
  } else {
    float p0 = 0.0f;
    rc = reduce_p0_seq_f32(s, q, &p0);
    if (rc) return rc;
    result = draw >= (double)p0;
    float pf = result ? 1.0f - p0 : p0;
    float nff = pf > 0.0f ? 1.0f / std::sqrt(pf) : 0.0f;
    prob = (double)pf;
    nf = (double)nff;
  }
  launch_collapse(s, q, result, nf);
  RQ_HIP(hipGetLastError());
  rq_measurement m{q, result, prob};
  s.record.push_back(m);
  if (out) *out = m;
  return RQ_OK;
}

int reset_impl(State& s, uint32_t q, double draw) {  // state.rs:344-352
  rq_measurement m;
  int rc = measure_impl(s, q, draw, &m);
  if (rc) return rc;
  if (m.result) {
    rq_gate x{};
    x.kind = RQ_GATE_X;
    x.q0 = q;
    rq_c64 mx[4];
    matrix_1q(x, mx);
    launch_apply_1q(s, q, mx);
    RQ_HIP(hipGetLastError());
  }
  return RQ_OK;
}

}  // namespace

int measure_state(State& s, uint32_t q, double draw, rq_measurement* out) { return measure_impl(s, q, draw, out); }

// mixed_precision.rs:260-278 (from_f64: narrowed amplitudes, the source's
// measurement record, gate count 0) and :280-283 (to_f64: widened amplitudes
// through QuantumState::from_amplitudes, so the qubit limit of the complex-f64
// state applies, and no measurement record).
int convert_state(const State& src, int to_prec, State** out) {
  uint32_t maxq = rq_max_qubits(to_prec);
  if (src.nq > maxq) return limit_error(src.nq, (int)maxq);
  State* d = nullptr;
  int rc = alloc_state(src.nq, to_prec, &d);
  if (rc) return rc;
  hipError_t e = hipStreamSynchronize(src.stream);
  if (e == hipSuccess) {
    launch_convert(*d, src);
    e = hipGetLastError();
  }
  if (e == hipSuccess) e = hipStreamSynchronize(d->stream);
  if (e != hipSuccess) {
    rq_state_destroy((rq_state*)d);
    return hip_fail(e, "precision conversion");
  }
  if (to_prec == RQ_C32) d->record = src.record;
  *out = d;
  return RQ_OK;
}

int reset_state(State& s, uint32_t q, double draw) { return reset_impl(s, q, draw); }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return RQ_ERR_DEVICE;
}

int ensure_red(State& s, size_t doubles) {
  if (doubles <= s.red_cap) return RQ_OK;
  if (s.red) {
    RQ_HIP(hipStreamSynchronize(s.stream));
    RQ_HIP(hipFree(s.red));
    s.red = nullptr;
    s.red_cap = 0;
  }
  size_t cap = doubles < 4096 ? 4096 : doubles;
  RQ_HIP(hipMalloc((void**)&s.red, cap * sizeof(double)));
  s.red_cap = cap;
  return RQ_OK;
}

// Bytes one exact kernel moves for a gate (reads + writes of the amplitudes
// it touches): dense 1q / Rzz / dense 2q 2Ns, CNOT and SWAP Ns, CZ Ns/2.
double gate_bytes(const State& s, int kind) {
  double ns = (double)s.n * (double)s.amp_bytes();
  switch (kind) {
    case RQ_GATE_CNOT:
    case RQ_GATE_SWAP: return ns;
    case RQ_GATE_CZ: return ns / 2;
    default: return 2 * ns;
  }
}

// state.rs:158-204: one gate, exact kernels.
int apply_gate_impl(State& s, const rq_gate& g, double draw, rq_measurement* out, int* has_out) {
  if (has_out) *has_out = 0;
  if (g.kind < RQ_GATE_H || g.kind > RQ_GATE_UNITARY1Q)
    return fail(RQ_ERR_CIRCUIT, "unsupported gate: kind " + std::to_string(g.kind));
  if (g.kind == RQ_GATE_BARRIER) return RQ_OK;
  int rc = check_qubit(s, g.q0);
  if (rc) return rc;
  if (is_2q(g.kind)) {
    rc = check_qubit(s, g.q1);
    if (rc) return rc;
  }
  switch (g.kind) {
    case RQ_GATE_MEASURE:
      rc = measure_impl(s, g.q0, draw, out);
      if (!rc && has_out) *has_out = 1;
      return rc;
    case RQ_GATE_RESET:
      return reset_impl(s, g.q0, draw);
    default:
      break;
  }
  if (is_2q(g.kind)) {
    if (g.q0 == g.q1)
      return fail(RQ_ERR_CIRCUIT, "two-qubit gate requires distinct qubits, got " + std::to_string(g.q0) + " and " +
                                      std::to_string(g.q1));
    switch (g.kind) {
      case RQ_GATE_CNOT: launch_cnot(s, g.q0, g.q1); break;
      case RQ_GATE_CZ: launch_cz(s, g.q0, g.q1); break;
      case RQ_GATE_SWAP: launch_swap(s, g.q0, g.q1); break;
      default: {
        rq_c64 m[16];
        matrix_2q(g, m);
        launch_diag_2q(s, g.q0, g.q1, m);
      }
    }
  } else {
    rq_c64 m[4];
    matrix_1q(g, m);
    launch_apply_1q(s, g.q0, m);
  }
  RQ_HIP(hipGetLastError());
  ++s.gate_count;
  return RQ_OK;
}

int run_gates(State& s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws,
              rq_measurement* meas, size_t cap, size_t* nout, int fusion, rq_metrics* metrics, bool reset) {
  if (ngates && !gates) return fail(RQ_ERR_ARGUMENT, "gates is null");
  auto t0 = std::chrono::steady_clock::now();
  hipEvent_t e0 = nullptr, e1 = nullptr;
  if (metrics) {
    RQ_HIP(hipEventCreate(&e0));
    RQ_HIP(hipEventCreate(&e1));
  }
  // A fused run from |0..0> lets its first stage make the amplitudes instead
  // of clearing the state first (run_fused clears it when no stage runs).
  const bool zero_first = reset && fusion == 1;
  if (reset) {
    if (!zero_first) launch_init_zero(s);
    s.record.clear();
    s.gate_count = 0;
  }
  if (e0) RQ_HIP(hipEventRecord(e0, s.stream));
  size_t cursor = 0, produced = 0;
  uint64_t unitary = 0, passes = 0, fused = 0;
  double hbm = 0;
  int rc = RQ_OK;
  if (fusion) {
    FusionStats st{};
    rc = fusion == 2 ? run_dense(s, gates, ngates, draws, ndraws, &cursor, meas, cap, &produced, &st)
                     : run_fused(s, gates, ngates, draws, ndraws, &cursor, meas, cap, &produced, &st, zero_first);
    unitary = st.unitary;
    passes = st.passes;
    fused = st.fused;
    hbm = st.hbm_bytes;
  } else {
    for (size_t k = 0; k < ngates && rc == RQ_OK; ++k) {
      const rq_gate& g = gates[k];
      double draw = 0.0;
      if (g.kind == RQ_GATE_MEASURE || g.kind == RQ_GATE_RESET) {
        if (cursor >= ndraws) {
          rc = fail(RQ_ERR_CIRCUIT, "measurement needs a uniform draw but the draw list is exhausted");
          break;
        }
        draw = draws[cursor++];
      }
      rq_measurement m;
      int has = 0;
      rc = apply_gate_impl(s, g, draw, &m, &has);
      if (rc) break;
      if (has) {
        if (meas && produced < cap) meas[produced] = m;
        ++produced;
      }
      if (!is_non_unitary(g.kind)) {
        ++unitary;
        ++passes;
        hbm += gate_bytes(s, g.kind);
      }
    }
  }
  if (nout) *nout = produced;
  if (rc == RQ_OK && e1) RQ_HIP(hipEventRecord(e1, s.stream));
  if (rc == RQ_OK) RQ_HIP(hipStreamSynchronize(s.stream));
  auto t1 = std::chrono::steady_clock::now();
  if (metrics && rc == RQ_OK) {
    float ms = 0.f;
    RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
    double secs = std::chrono::duration<double>(t1 - t0).count();
    metrics->num_qubits = s.nq;
    metrics->gate_count = unitary;
    metrics->execution_time_ns = (uint64_t)(secs * 1e9);
    metrics->peak_memory_bytes = s.bytes();
    metrics->gates_per_second = secs > 0 ? unitary / secs : 0.0;
    metrics->gates_fused = fused;
    metrics->passes = passes;
    metrics->device_time_ms = ms;
    metrics->hbm_bytes = hbm;
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  return rc;
}

}  // namespace rq

using namespace rq;

extern "C" {

const char* rq_last_error(void) { return g_last_error.c_str(); }

uint32_t rq_max_qubits(int precision) {
  compute_limits();
  return (uint32_t)(precision == RQ_C32 ? g_max_q32 : g_max_q64);
}

int rq_device_synchronize(void) {
  int rc = device_ready();
  if (rc) return rc;
  RQ_HIP(hipDeviceSynchronize());
  return RQ_OK;
}

int rq_state_create(uint32_t nq, int precision, int with_seed, rq_state** out) {
  if (!out) return fail(RQ_ERR_ARGUMENT, "out is null");
  if (precision != RQ_C64 && precision != RQ_C32) return fail(RQ_ERR_ARGUMENT, "unknown precision");
  // state.rs:32-36 (new(0) errors; new_with_seed(0) is accepted for f64),
  // mixed_precision.rs:209-213,234-238 (f32 rejects 0 in both).
  if (nq == 0 && (precision == RQ_C32 || !with_seed))
    return fail(RQ_ERR_CIRCUIT, "cannot create quantum state with 0 qubits");
  uint32_t maxq = rq_max_qubits(precision);
  if (nq > maxq) return limit_error(nq, (int)maxq);
  State* s = nullptr;
  int rc = alloc_state(nq, precision, &s);
  if (rc) return rc;
  launch_init_zero(*s);
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  if (e != hipSuccess) {
    rq_state_destroy((rq_state*)s);
    return hip_fail(e, "init");
  }
  *out = (rq_state*)s;
  return RQ_OK;
}

int rq_state_from_amplitudes(const void* host, size_t len, uint32_t nq, int precision, rq_state** out) {
  if (!out || (!host && len)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (precision != RQ_C64 && precision != RQ_C32) return fail(RQ_ERR_ARGUMENT, "unknown precision");
  uint32_t maxq = rq_max_qubits(precision);
  if (nq > maxq) return limit_error(nq, (int)maxq);  // state.rs:77-82
  if (len != ((size_t)1 << nq))                      // state.rs:83-89
    return fail(RQ_ERR_INVALID_STATE, "invalid state vector: length " + std::to_string(len) +
                                          " does not match 2^" + std::to_string(nq));
  State* s = nullptr;
  int rc = alloc_state(nq, precision, &s);
  if (rc) return rc;
  hipError_t e = hipMemcpy(s->amps, host, s->bytes(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rq_state_destroy((rq_state*)s);
    return hip_fail(e, "from_amplitudes");
  }
  *out = (rq_state*)s;
  return RQ_OK;
}

size_t rq_state_checkpoint_bytes(const rq_state* h) {  // replay.rs:290-293
  const State* s = (const State*)h;
  return s ? s->n * 16 : 0;
}

// replay.rs:257-268: interleaved little-endian f64 (re, im).  A complex-f64
// state's device buffer is exactly that; complex-f32 amplitudes are copied
// out in chunks and widened (exact).
int rq_state_checkpoint(const rq_state* h, void* out) {
  const State* s = (const State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  RQ_HIP(hipStreamSynchronize(s->stream));
  if (s->prec == RQ_C64) {
    RQ_HIP(hipMemcpy(out, s->amps, s->n * 16, hipMemcpyDeviceToHost));
    return RQ_OK;
  }
  const size_t chunk = size_t(1) << 22;  // amplitudes per staging copy
  std::vector<float> stage(2 * std::min(chunk, s->n));
  double* o = (double*)out;
  for (size_t i0 = 0; i0 < s->n; i0 += chunk) {
    const size_t m = std::min(chunk, s->n - i0);
    RQ_HIP(hipMemcpy(stage.data(), (const float*)s->amps + 2 * i0, m * 8, hipMemcpyDeviceToHost));
    for (size_t k = 0; k < 2 * m; ++k) o[2 * i0 + k] = (double)stage[k];
  }
  return RQ_OK;
}

// replay.rs:270-287 then QuantumState::from_amplitudes (state.rs:77-96).
int rq_state_from_checkpoint(const void* bytes, size_t nbytes, int precision, rq_state** out) {
  if (!out || (!bytes && nbytes)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (precision != RQ_C64 && precision != RQ_C32) return fail(RQ_ERR_ARGUMENT, "unknown precision");
  const size_t count = nbytes / 16;
  if (nbytes % 16 || count == 0 || (count & (count - 1)))
    return fail(RQ_ERR_INVALID_STATE, "invalid state vector: checkpoint of " + std::to_string(nbytes) +
                                          " bytes is not 16 * 2^n");
  uint32_t nq = 0;
  while ((size_t(1) << nq) < count) ++nq;
  if (precision == RQ_C64) return rq_state_from_amplitudes(bytes, count, nq, RQ_C64, out);
  uint32_t maxq = rq_max_qubits(precision);
  if (nq > maxq) return limit_error(nq, (int)maxq);
  State* s = nullptr;
  int rc = alloc_state(nq, precision, &s);
  if (rc) return rc;
  const size_t chunk = size_t(1) << 22;
  std::vector<float> stage(2 * std::min(chunk, count));
  const double* in = (const double*)bytes;
  for (size_t i0 = 0; i0 < count; i0 += chunk) {
    const size_t m = std::min(chunk, count - i0);
    for (size_t k = 0; k < 2 * m; ++k) stage[k] = (float)in[2 * i0 + k];
    hipError_t e = hipMemcpy((float*)s->amps + 2 * i0, stage.data(), m * 8, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
      rq_state_destroy((rq_state*)s);
      return hip_fail(e, "from_checkpoint");
    }
  }
  *out = (rq_state*)s;
  return RQ_OK;
}

void rq_state_destroy(rq_state* h) {
  State* s = (State*)h;
  if (!s) return;
  if (s->stream) (void)hipStreamSynchronize(s->stream);
  if (s->amps) (void)hipFree(s->amps);
  if (s->red) (void)hipFree(s->red);
  if (s->samp) (void)hipFree(s->samp);
  if (s->prog) (void)hipFree(s->prog);
  release_fused_program(*s);
  if (s->host_red) (void)hipHostFree(s->host_red);
  if (s->stream) (void)hipStreamDestroy(s->stream);
  delete s;
}

uint32_t rq_state_num_qubits(const rq_state* h) { return h ? ((const State*)h)->nq : 0; }
int rq_state_precision(const rq_state* h) { return h ? ((const State*)h)->prec : -1; }

size_t rq_estimate_memory(uint32_t nq, int precision) {  // state.rs:144-146
  return ((size_t)1 << nq) * (precision == RQ_C32 ? 8 : 16);
}

}  // extern "C"

namespace {
// Reads of states up to 64 MiB go through one pinned host buffer that the
// device copy kernel fills on the state's stream, then a host memcpy.  A
// device-to-host copy through the runtime's copy engine paid 2-14 ms on the
// first read of each newly allocated state (about 1 ms once warm) for 16 MB;
// the kernel path takes no such start-up.  The buffer grows to the largest
// staged read and is kept until the process ends; reads from several threads
// take turns on it.
constexpr size_t STAGED_READ_MAX = size_t(64) << 20;
std::mutex g_stage_mu;
void* g_stage = nullptr;
size_t g_stage_cap = 0;
}  // namespace

extern "C" {

int rq_state_read(const rq_state* h, void* host) {
  const State* s = (const State*)h;
  if (!s || !host) return fail(RQ_ERR_ARGUMENT, "null pointer");
  const size_t bytes = s->bytes();
  if (bytes <= STAGED_READ_MAX && bytes % 16 == 0) {
    std::lock_guard<std::mutex> lock(g_stage_mu);
    if (g_stage_cap < bytes) {
      void* p = nullptr;
      if (hipHostMalloc(&p, STAGED_READ_MAX, hipHostMallocDefault) == hipSuccess) {
        if (g_stage) (void)hipHostFree(g_stage);
        g_stage = p;
        g_stage_cap = STAGED_READ_MAX;
      } else {
        (void)hipGetLastError();  // no pinned memory: the copy-engine path below
      }
    }
    if (g_stage_cap >= bytes) {
      launch_plain_copy(g_stage, s->amps, bytes, s->stream);
      RQ_HIP(hipGetLastError());
      RQ_HIP(hipStreamSynchronize(s->stream));
      std::memcpy(host, g_stage, bytes);
      return RQ_OK;
    }
  }
  RQ_HIP(hipMemcpyAsync(host, s->amps, bytes, hipMemcpyDeviceToHost, s->stream));
  RQ_HIP(hipStreamSynchronize(s->stream));
  return RQ_OK;
}

int rq_state_write(rq_state* h, const void* host) {
  State* s = (State*)h;
  if (!s || !host) return fail(RQ_ERR_ARGUMENT, "null pointer");
  RQ_HIP(hipMemcpyAsync(s->amps, host, s->bytes(), hipMemcpyHostToDevice, s->stream));
  RQ_HIP(hipStreamSynchronize(s->stream));
  return RQ_OK;
}

int rq_state_set_zero(rq_state* h) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  launch_init_zero(*s);
  RQ_HIP(hipGetLastError());
  s->record.clear();
  s->gate_count = 0;
  return RQ_OK;
}

int rq_state_copy(const rq_state* hs, rq_state* hd) {
  const State* src = (const State*)hs;
  State* dst = (State*)hd;
  if (!src || !dst) return fail(RQ_ERR_ARGUMENT, "null state");
  if (src->nq != dst->nq || src->prec != dst->prec) return fail(RQ_ERR_ARGUMENT, "state shapes differ");
  RQ_HIP(hipStreamSynchronize(src->stream));
  launch_copy(*dst, *src);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipStreamSynchronize(dst->stream));
  return RQ_OK;
}

int rq_state_from_f64(const rq_state* h, rq_state** out) {
  const State* s = (const State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (s->prec != RQ_C64) return fail(RQ_ERR_ARGUMENT, "from_f64 needs a complex-f64 state");
  State* d = nullptr;
  int rc = convert_state(*s, RQ_C32, &d);
  if (rc) return rc;
  *out = (rq_state*)d;
  return RQ_OK;
}

int rq_state_to_f64(const rq_state* h, rq_state** out) {
  const State* s = (const State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (s->prec != RQ_C32) return fail(RQ_ERR_ARGUMENT, "to_f64 needs a complex-f32 state");
  State* d = nullptr;
  int rc = convert_state(*s, RQ_C64, &d);
  if (rc) return rc;
  *out = (rq_state*)d;
  return RQ_OK;
}

double rq_precision_error_bound(const rq_state* h) {  // mixed_precision.rs:325-327
  const State* s = (const State*)h;
  if (!s || s->prec != RQ_C32) return 0.0;
  return (double)s->gate_count * (double)1.1920929e-07f;
}

int rq_apply_gate(rq_state* h, const rq_gate* g, double draw, rq_measurement* out, int* has_out) {
  State* s = (State*)h;
  if (!s || !g) return fail(RQ_ERR_ARGUMENT, "null pointer");
  return apply_gate_impl(*s, *g, draw, out, has_out);
}

int rq_apply_1q(rq_state* h, uint32_t q, const rq_c64* m) {
  State* s = (State*)h;
  if (!s || !m) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = check_qubit(*s, q);
  if (rc) return rc;
  launch_apply_1q(*s, q, m);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

int rq_apply_2q(rq_state* h, uint32_t q1, uint32_t q2, const rq_c64* m) {
  State* s = (State*)h;
  if (!s || !m) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = check_qubit(*s, q1);
  if (!rc) rc = check_qubit(*s, q2);
  if (rc) return rc;
  if (q1 == q2)
    return fail(RQ_ERR_CIRCUIT, "two-qubit gate requires distinct qubits, got " + std::to_string(q1) + " and " +
                                    std::to_string(q2));
  launch_apply_2q(*s, q1, q2, m);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

int rq_run_circuit(rq_state* h, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws,
                   rq_measurement* meas, size_t cap, size_t* nout, int fusion, rq_metrics* metrics) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  return run_gates(*s, gates, ngates, draws, ndraws, meas, cap, nout, fusion, metrics, true);
}

// simulator.rs:56-101 with a noise model.  The channel's choices depend on
// the draws only, never on the state, so the circuit is expanded on the host
// into gates plus the Paulis the channel applies (apply_noise,
// simulator.rs:178-221, per qubit of Gate::qubits, gate.rs:40-64) and run
// through the ordinary path; the Measure / Reset draws are taken from the
// same stream in the reference's order (the gate's draw before its noise).
int rq_run_circuit_noisy(rq_state* h, const rq_gate* gates, size_t ngates, const rq_noise* noise,
                         const double* draws, size_t ndraws, size_t* draws_used, rq_measurement* meas,
                         size_t cap, size_t* nout, int fusion, rq_metrics* metrics) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  if ((ngates && !gates) || (ndraws && !draws)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  const double dp = noise ? noise->depolarizing_rate : 0.0, bf = noise ? noise->bit_flip_rate : 0.0,
               pf = noise ? noise->phase_flip_rate : 0.0;
  std::vector<rq_gate> ex;
  std::vector<double> md;
  ex.reserve(ngates + ngates / 4 + 8);
  size_t cur = 0;
  uint64_t inserted = 0;
  const std::string exhausted = "noise channel needs a uniform draw but the draw list is exhausted";
  auto pauli = [&](int kind, uint32_t q) {
    rq_gate p{};
    p.kind = kind;
    p.q0 = q;
    ex.push_back(p);
    ++inserted;
  };
  for (size_t k = 0; k < ngates; ++k) {
    const rq_gate& g = gates[k];
    if (g.kind == RQ_GATE_MEASURE || g.kind == RQ_GATE_RESET) {
      if (cur >= ndraws) return fail(RQ_ERR_CIRCUIT, "measurement needs a uniform draw but the draw list is exhausted");
      md.push_back(draws[cur++]);
    }
    ex.push_back(g);
    const uint32_t qs[2] = {g.q0, g.q1};
    const int nqs = g.kind == RQ_GATE_BARRIER ? 0 : (is_2q(g.kind) ? 2 : 1);
    for (int t = 0; t < nqs; ++t) {
      if (dp > 0.0) {
        if (cur >= ndraws) return fail(RQ_ERR_CIRCUIT, exhausted);
        if (draws[cur++] < dp) {
          if (cur >= ndraws) return fail(RQ_ERR_CIRCUIT, exhausted);
          const double choice = draws[cur++];
          pauli(choice < 1.0 / 3.0 ? RQ_GATE_X : (choice < 2.0 / 3.0 ? RQ_GATE_Y : RQ_GATE_Z), qs[t]);
        }
      }
      if (bf > 0.0) {
        if (cur >= ndraws) return fail(RQ_ERR_CIRCUIT, exhausted);
        if (draws[cur++] < bf) pauli(RQ_GATE_X, qs[t]);
      }
      if (pf > 0.0) {
        if (cur >= ndraws) return fail(RQ_ERR_CIRCUIT, exhausted);
        if (draws[cur++] < pf) pauli(RQ_GATE_Z, qs[t]);
      }
    }
  }
  if (draws_used) *draws_used = cur;
  int rc = run_gates(*s, ex.data(), ex.size(), md.data(), md.size(), meas, cap, nout, fusion, metrics, true);
  if (rc == RQ_OK && metrics) {  // the channel's Paulis are not circuit gates (simulator.rs:73-75)
    metrics->gate_count -= inserted;
    const double secs = metrics->execution_time_ns * 1e-9;
    metrics->gates_per_second = secs > 0 ? metrics->gate_count / secs : 0.0;
  }
  return rc;
}

int rq_apply_circuit(rq_state* h, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws,
                     rq_measurement* meas, size_t cap, size_t* nout, int fusion, rq_metrics* metrics) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  return run_gates(*s, gates, ngates, draws, ndraws, meas, cap, nout, fusion, metrics, false);
}

uint32_t rq_circuit_depth(uint32_t nq, const rq_gate* gates, size_t ngates) {  // circuit.rs:153-184
  std::vector<uint32_t> d(nq, 0);
  for (size_t k = 0; k < ngates; ++k) {
    const rq_gate& g = gates[k];
    if (g.kind == RQ_GATE_BARRIER) {
      uint32_t mx = 0;
      for (uint32_t v : d) mx = v > mx ? v : mx;
      for (auto& v : d) v = mx;
      continue;
    }
    uint32_t qs[2] = {g.q0, g.q1};
    int c = is_2q(g.kind) ? 2 : 1;
    uint32_t mx = 0;
    for (int t = 0; t < c; ++t) {
      uint32_t v = qs[t] < nq ? d[qs[t]] : 0;
      mx = v > mx ? v : mx;
    }
    for (int t = 0; t < c; ++t)
      if (qs[t] < nq) d[qs[t]] = mx + 1;
  }
  uint32_t mx = 0;
  for (uint32_t v : d) mx = v > mx ? v : mx;
  return mx;
}

int rq_measure(rq_state* h, uint32_t q, double draw, rq_measurement* out) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  return measure_impl(*s, q, draw, out);
}

int rq_measure_all(rq_state* h, const double* draws, rq_measurement* out) {
  State* s = (State*)h;
  if (!s || !draws) return fail(RQ_ERR_ARGUMENT, "null pointer");
  for (uint32_t q = 0; q < s->nq; ++q) {
    int rc = measure_impl(*s, q, draws[q], out ? &out[q] : nullptr);
    if (rc) return rc;
  }
  return RQ_OK;
}

int rq_reset_qubit(rq_state* h, uint32_t q, double draw) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  return reset_impl(*s, q, draw);
}

size_t rq_measurement_record(const rq_state* h, rq_measurement* out, size_t cap) {
  const State* s = (const State*)h;
  if (!s) return 0;
  for (size_t k = 0; k < s->record.size() && k < cap && out; ++k) out[k] = s->record[k];
  return s->record.size();
}

int rq_sample_shots(const rq_state* h, const double* draws, size_t shots, uint8_t* bits, uint64_t* near_ties) {
  State* s = (State*)h;
  if (!s || (shots && (!draws || !bits))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  return sample_shots(*s, draws, shots, bits, near_ties);
}

int rq_probabilities(const rq_state* h, double* out) {
  State* s = (State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  double* d = nullptr;
  RQ_HIP(hipMalloc((void**)&d, s->n * sizeof(double)));
  int rc = launch_probabilities(*s, d);
  hipError_t e = hipSuccess;
  if (!rc) e = hipMemcpyAsync(out, d, s->n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(s->stream);
  (void)hipFree(d);
  if (!rc && e != hipSuccess) rc = hip_fail(e, "probabilities");
  return rc;
}

int rq_probability_of_qubit(const rq_state* h, uint32_t q, double* out) {  // state.rs:128-137
  State* s = (State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (q >= 64) return check_qubit(*s, q);
  if (q >= s->nq) {  // the reference's loop finds no index with that bit set
    *out = 0.0;
    return RQ_OK;
  }
  return reduce_masked_normsq(*s, 1ull << q, 1ull << q, out);
}

}  // extern "C"

namespace rq {
// PauliString (types.rs:201-203) -> (x mask, Y/Z parity mask, #Y).  Factors
// apply in order exactly as state.rs:370-392: X/Y flip bit q of j, Y/Z read
// bit q of i; a repeated qubit composes the same way.  Z or I on a qubit at or
// beyond num_qubits reads a zero bit (no effect) as in the reference; X or Y
// there would index out of range (a panic in the reference) and is an error.
int canon_pauli(const State& s, const rq_pauli_factor* ops, size_t nops, PauliCanon* out) {
  PauliCanon c{0, 0, 0};
  for (size_t k = 0; k < nops; ++k) {
    uint32_t q = ops[k].qubit;
    int op = ops[k].op;
    if (op < 0 || op > 3) return fail(RQ_ERR_ARGUMENT, "unknown Pauli operator");
    if (op == RQ_PAULI_I) continue;
    if (q >= s.nq && (op == RQ_PAULI_X || op == RQ_PAULI_Y || q >= 64)) return check_qubit(s, q);
    if (q >= s.nq) continue;
    uint64_t b = 1ull << q;
    if (op == RQ_PAULI_X || op == RQ_PAULI_Y) c.x ^= b;
    if (op == RQ_PAULI_Y || op == RQ_PAULI_Z) c.zy ^= b;
    if (op == RQ_PAULI_Y) c.ny += 1;
  }
  c.ny &= 3;
  *out = c;
  return RQ_OK;
}
}  // namespace rq

extern "C" {

int rq_expectation(const rq_state* h, const rq_pauli_factor* ops, size_t nops, double* out) {
  State* s = (State*)h;
  if (!s || !out || (nops && !ops)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  PauliCanon c;
  int rc = canon_pauli(*s, ops, nops, &c);
  if (rc) return rc;
  return reduce_pauli(*s, &c, 1, out);
}

int rq_expectation_hamiltonian(const rq_state* h, const double* coeffs, const uint64_t* offsets,
                               const rq_pauli_factor* ops, size_t nterms, double* out) {
  State* s = (State*)h;
  if (!s || !out || (nterms && (!coeffs || !offsets))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  std::vector<PauliCanon> cs(nterms);
  for (size_t k = 0; k < nterms; ++k) {
    if (offsets[k + 1] < offsets[k]) return fail(RQ_ERR_ARGUMENT, "offsets must be non-decreasing");
    int rc = canon_pauli(*s, ops + offsets[k], offsets[k + 1] - offsets[k], &cs[k]);
    if (rc) return rc;
  }
  std::vector<double> ev(nterms);
  if (nterms) {
    int rc = reduce_pauli(*s, cs.data(), nterms, ev.data());
    if (rc) return rc;
  }
  double acc = 0.0;  // state.rs:403-408: sequential sum of c_k * <P_k>
  for (size_t k = 0; k < nterms; ++k) acc += coeffs[k] * ev[k];
  *out = acc;
  return RQ_OK;
}

int rq_expectation_hamiltonian_noalloc(const rq_state* h, const double* coeffs, const uint64_t* offsets,
                                       const rq_pauli_factor* ops, size_t nterms, double* out) {
  State* s = (State*)h;
  if (!s || !out || (nterms && (!coeffs || !offsets))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  constexpr size_t BATCH = 32;  // terms per launch (the scratch holds one launch)
  PauliCanon cs[BATCH];
  double acc = 0.0;  // state.rs:403-408: sequential sum of c_k * <P_k>
  for (size_t t0 = 0; t0 < nterms; t0 += BATCH) {
    const size_t cnt = nterms - t0 < BATCH ? nterms - t0 : BATCH;
    for (size_t k = 0; k < cnt; ++k) {
      const size_t t = t0 + k;
      if (offsets[t + 1] < offsets[t]) return fail(RQ_ERR_ARGUMENT, "offsets must be non-decreasing");
      int rc = canon_pauli(*s, ops + offsets[t], offsets[t + 1] - offsets[t], &cs[k]);
      if (rc) return rc;
    }
    int rc = reduce_hamiltonian_noalloc(*s, coeffs + t0, cs, cnt, &acc);
    if (rc) return rc;
  }
  *out = acc;
  return RQ_OK;
}

int rq_expectation_z_all(const rq_state* h, double* out) {
  State* s = (State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  return reduce_z_all(*s, out);
}

int rq_norm_sq(const rq_state* h, double* out) {
  State* s = (State*)h;
  if (!s || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  return reduce_masked_normsq(*s, 0, 0, out);
}

int rq_normalize(rq_state* h) {  // state.rs:415-423
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  double ns = 0.0;
  int rc = reduce_masked_normsq(*s, 0, 0, &ns);
  if (rc) return rc;
  if (ns > 0.0) {
    launch_scale(*s, 1.0 / std::sqrt(ns));
    RQ_HIP(hipGetLastError());
  }
  return RQ_OK;
}

int rq_fidelity(const rq_state* ha, const rq_state* hb, double* out) {  // state.rs:426-435
  const State* a = (const State*)ha;
  const State* b = (const State*)hb;
  if (!a || !b || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (a->nq != b->nq) {
    *out = 0.0;
    return RQ_OK;
  }
  if (a->prec != b->prec) return fail(RQ_ERR_ARGUMENT, "fidelity needs states of one precision");
  double re = 0, im = 0;
  int rc = reduce_inner(*a, *b, &re, &im);
  if (rc) return rc;
  *out = re * re + im * im;
  return RQ_OK;
}

int rq_time_1q(rq_state* h, uint32_t q, const rq_c64* m, int reps, double* avg_ms) {
  State* s = (State*)h;
  if (!s || !m || !avg_ms || reps <= 0) return fail(RQ_ERR_ARGUMENT, "bad argument");
  int rc = check_qubit(*s, q);
  if (rc) return rc;
  hipEvent_t e0, e1;
  RQ_HIP(hipEventCreate(&e0));
  RQ_HIP(hipEventCreate(&e1));
  RQ_HIP(hipEventRecord(e0, s->stream));
  for (int r = 0; r < reps; ++r) launch_apply_1q(*s, q, m);
  RQ_HIP(hipEventRecord(e1, s->stream));
  RQ_HIP(hipEventSynchronize(e1));
  float ms = 0;
  RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  *avg_ms = ms / reps;
  return RQ_OK;
}

int rq_time_copy(rq_state* h, int reps, double* avg_ms) {
  State* s = (State*)h;
  if (!s || !avg_ms || reps <= 0) return fail(RQ_ERR_ARGUMENT, "bad argument");
  void* dst = nullptr;
  hipError_t e = hipMalloc(&dst, s->bytes());
  if (e != hipSuccess) {
    (void)hipGetLastError();
    return fail(RQ_ERR_ALLOC, "copy buffer allocation failed");
  }
  hipEvent_t e0, e1;
  RQ_HIP(hipEventCreate(&e0));
  RQ_HIP(hipEventCreate(&e1));
  launch_plain_copy(dst, s->amps, s->bytes(), s->stream);  // warm-up
  RQ_HIP(hipEventRecord(e0, s->stream));
  for (int r = 0; r < reps; ++r) launch_plain_copy(dst, s->amps, s->bytes(), s->stream);
  RQ_HIP(hipEventRecord(e1, s->stream));
  RQ_HIP(hipEventSynchronize(e1));
  float ms = 0;
  RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipFree(dst);
  *avg_ms = ms / reps;
  return RQ_OK;
}

}  // extern "C"

extern "C" {

// Tuning knobs (no reference counterpart).  key 0: streaming unroll (1, 2, 4,
// 8); key 1: non-temporal loads/stores (0/1); key 2: fused tile qubits m;
// key 3: fused low qubits c; key 4: fused matrix-core threshold (VALU
// instructions per wave of a complex-f32 phase's ops; -1: never); key 5:
// tile-qubit seeds the fused planner tries per stage (0: plain greedy scan).
int rq_tune_set(int key, int value) {
  switch (key) {
    case 0:
      if (value != 1 && value != 2 && value != 4 && value != 8) return fail(RQ_ERR_ARGUMENT, "unroll must be 1, 2, 4 or 8");
      stream_tune().unroll = value;
      return RQ_OK;
    case 1:
      stream_tune().nontemporal = value != 0;
      return RQ_OK;
    case 2:
      if (value < 4 || value > FUSE_MAX_M) return fail(RQ_ERR_ARGUMENT, "tile qubits out of range");
      fusion_config().m = value;
      return RQ_OK;
    case 3:
      if (value < 0 || value > FUSE_MAX_M) return fail(RQ_ERR_ARGUMENT, "low qubits out of range");
      fusion_config().c = value;
      return RQ_OK;
    case 4:
      if (value < -1 || value > (1 << 20)) return fail(RQ_ERR_ARGUMENT, "matrix-core threshold out of range");
      fusion_config().mf = value;
      return RQ_OK;
    case 5:
      if (value < 0 || value > 4096) return fail(RQ_ERR_ARGUMENT, "planner seeds out of range");
      fusion_config().seeds = value;
      return RQ_OK;
    case 6:
      if (value < -1 || value > 65536 || (value > 0 && (value & (value - 1))))
        return fail(RQ_ERR_ARGUMENT, "block chunk must be -1, 0 or a power of two up to 65536");
      stream_tune().xcd_chunk = value;
      return RQ_OK;
    default:
      return fail(RQ_ERR_ARGUMENT, "unknown tuning key");
  }
}

int rq_tune_get(int key) {
  switch (key) {
    case 0: return stream_tune().unroll;
    case 1: return stream_tune().nontemporal ? 1 : 0;
    case 2: return fusion_config().m;
    case 3: return fusion_config().c;
    case 4: return fusion_config().mf;
    case 5: return fusion_config().seeds;
    case 6: return stream_tune().xcd_chunk;
    default: return -2;
  }
}

}  // extern "C"
