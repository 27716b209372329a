// ORACLE — test infrastructure only.
//
// A CPU restatement of ruqu-core's scalar state-vector path and of the
// ruqu-algorithms QAOA/VQE drivers, written op-for-op after the Rust sources so
// that, built with -ffp-contract=off and glibc sin/cos, its f64 results follow
// the same IEEE operation sequence as the reference.  Only tests/, the smoke
// check and the CPU-baseline leg of bench.py may load this library; the product
// (libruqu_gpu.so) never links or calls it.
//
// Every function cites the reference file:line it follows (paths relative to
// the ruvnet/ruqu checkout).  How each one is pinned (known-answer tests from
// the reference's own test files, closed forms, brute-force Kronecker products,
// invariants) is listed in tests/test_oracle_*.py.
//
// RNG: the reference draws uniforms from rand 0.8 StdRng (ChaCha12), which is
// not vendored; every draw is therefore an explicit input here ("fed the
// reference's uniform draws").  The seeded stream itself is restated in
// stdrng_oracle.py and pinned there by the RFC 7539 and PCG32 vectors.

#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <vector>

// Mutation testing of the pinning tests.  ORC_PLANT = 0 (the default, and the
// only value the Makefile builds) is the faithful restatement.  A test build
// with ORC_PLANT = k plants one deliberate mistake in the optimizer loops or
// the ansatz builder, so the pinning tests can show that each of their checks
// fails on a plausible error:
//   1 run_qaoa: gamma step with the wrong sign
//   2 run_qaoa: gamma gradients all taken at the old gammas (not in sequence)
//   3 run_qaoa: shift applied to the next parameter instead of parameter i
//   4 run_qaoa: argmax keeps the first maximal index instead of the last
//   5 run_qaoa: convergence measured against the first cost, not the previous
//   6 run_qaoa: best bitstring taken after the update (at the new angles)
//   7 run_vqe: descent step with the wrong sign
//   8 run_vqe: parameters updated one by one (gradient not simultaneous)
//   9 run_vqe: best parameters tracked at the maximum energy
//  10 build_ansatz: CNOT chain reversed (q+1 -> q)
//  11 build_ansatz: Rz layer before the Ry layer
//  12 build_ansatz: parameter index restarts at 0 in every layer
#ifndef ORC_PLANT
#define ORC_PLANT 0
#endif

namespace {

// ---------------------------------------------------------------------------
// types.rs:8-90 — Complex (f64), no FMA: mul = (a.re*b.re - a.im*b.im,
// a.re*b.im + a.im*b.re).
// ---------------------------------------------------------------------------
struct C64 {
  double re, im;
};
inline C64 c64(double re, double im) { return C64{re, im}; }
inline C64 operator+(C64 a, C64 b) { return C64{a.re + b.re, a.im + b.im}; }
inline C64 operator-(C64 a) { return C64{-a.re, -a.im}; }
inline C64 operator*(C64 a, C64 b) {
  return C64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline C64 scale(C64 a, double s) { return C64{a.re * s, a.im * s}; }  // types.rs:129-138
inline C64 conj(C64 a) { return C64{a.re, -a.im}; }                     // types.rs:42-47
inline double norm_sq(C64 a) { return a.re * a.re + a.im * a.im; }     // types.rs:32-34
// (cos th, sin th) of one argument through one glibc sincos call.  Rust's
// f64::cos/sin lower to LLVM intrinsics, and LLVM turns a cos/sin pair of one
// argument into a single `sincos` libcall on GNU targets; glibc's sincos and
// its separate sin/cos can differ in the last bit, so both the oracle and the
// engine call sincos explicitly wherever the reference evaluates such a pair.
inline void cos_sin(double th, double* c, double* s) { sincos(th, s, c); }
inline C64 from_polar(double r, double th) {                            // types.rs:24-29
  double c, s;
  cos_sin(th, &c, &s);
  return C64{r * c, r * s};
}

// mixed_precision.rs:33-162 — Complex32, same op order in f32.
struct C32 {
  float re, im;
};
inline C32 operator+(C32 a, C32 b) { return C32{a.re + b.re, a.im + b.im}; }
inline C32 operator*(C32 a, C32 b) {
  return C32{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re};
}
inline C32 scale32(C32 a, float s) { return C32{a.re * s, a.im * s}; }
inline float norm_sq32(C32 a) { return a.re * a.re + a.im * a.im; }
inline C32 from_f64(C64 c) { return C32{(float)c.re, (float)c.im}; }  // mixed_precision.rs:80-85

const double FRAC_1_SQRT_2 = 0.70710678118654752440084436210484903928483593768847;

}  // namespace

extern "C" {

// Gate record shared with the tests: the 20 variants of gate.rs:8-36 in
// declaration order.  Layout: kind, q0, q1, pad, theta, u[2][2] (re,im pairs).
typedef struct {
  int32_t kind;
  uint32_t q0, q1, pad;
  double theta;
  double u[8];
} orc_gate;

enum {
  G_H = 0, G_X, G_Y, G_Z, G_S, G_SDG, G_T, G_TDG, G_RX, G_RY, G_RZ, G_PHASE,
  G_CNOT, G_CZ, G_SWAP, G_RZZ, G_MEASURE, G_RESET, G_BARRIER, G_UNITARY1Q
};

// error.rs:8-23 variants in declaration order.
enum { E_OK = 0, E_QUBIT_LIMIT = 1, E_INVALID_QUBIT = 2, E_ALLOC = 3, E_INVALID_STATE = 4, E_CIRCUIT = 5 };

typedef struct {
  uint32_t qubit;
  int32_t result;
  double probability;
} orc_outcome;

}  // extern "C"

namespace {

bool is_1q_kind(int k) { return (k >= G_H && k <= G_PHASE) || k == G_UNITARY1Q; }
bool is_2q_kind(int k) { return k >= G_CNOT && k <= G_RZZ; }
bool is_non_unitary(int k) { return k == G_MEASURE || k == G_RESET || k == G_BARRIER; }  // gate.rs:67-69

// gate.rs:72-148 — matrix_1q.  Returns false for non-1q gates.
bool matrix_1q(const orc_gate& g, C64 m[2][2]) {
  const C64 c0{0.0, 0.0}, c1{1.0, 0.0}, ci{0.0, 1.0};
  switch (g.kind) {
    case G_H: {  // gate.rs:79-82
      C64 h{FRAC_1_SQRT_2, 0.0};
      m[0][0] = h; m[0][1] = h; m[1][0] = h; m[1][1] = -h;
      return true;
    }
    case G_X: m[0][0] = c0; m[0][1] = c1; m[1][0] = c1; m[1][1] = c0; return true;   // :85
    case G_Y: m[0][0] = c0; m[0][1] = -ci; m[1][0] = ci; m[1][1] = c0; return true;  // :88
    case G_Z: m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = -c1; return true;  // :91
    case G_S: m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = ci; return true;   // :94
    case G_SDG: m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = -ci; return true;  // :97
    case G_T: {  // :100-103
      C64 t{FRAC_1_SQRT_2, FRAC_1_SQRT_2};
      m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = t;
      return true;
    }
    case G_TDG: {  // :106-109
      C64 t{FRAC_1_SQRT_2, -FRAC_1_SQRT_2};
      m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = t;
      return true;
    }
    case G_RX: {  // :112-117
      double half = g.theta / 2.0, ch, sh;
      cos_sin(half, &ch, &sh);
      C64 c{ch, 0.0};
      C64 s{0.0, -sh};
      m[0][0] = c; m[0][1] = s; m[1][0] = s; m[1][1] = c;
      return true;
    }
    case G_RY: {  // :120-128
      double half = g.theta / 2.0, ch, sh;
      cos_sin(half, &ch, &sh);
      m[0][0] = C64{ch, 0.0}; m[0][1] = C64{-sh, 0.0};
      m[1][0] = C64{sh, 0.0}; m[1][1] = C64{ch, 0.0};
      return true;
    }
    case G_RZ: {  // :131-137
      double half = g.theta / 2.0;
      m[0][0] = from_polar(1.0, -half); m[0][1] = c0;
      m[1][0] = c0; m[1][1] = from_polar(1.0, half);
      return true;
    }
    case G_PHASE:  // :140
      m[0][0] = c1; m[0][1] = c0; m[1][0] = c0; m[1][1] = from_polar(1.0, g.theta);
      return true;
    case G_UNITARY1Q:  // :143
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) m[r][c] = C64{g.u[(r * 2 + c) * 2], g.u[(r * 2 + c) * 2 + 1]};
      return true;
    default:
      return false;
  }
}

// gate.rs:154-199 — matrix_2q, index = q1_bit*2 + q2_bit.
bool matrix_2q(const orc_gate& g, C64 m[4][4]) {
  const C64 c0{0.0, 0.0}, c1{1.0, 0.0};
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) m[r][c] = c0;
  switch (g.kind) {
    case G_CNOT:  // :161-166
      m[0][0] = c1; m[1][1] = c1; m[2][3] = c1; m[3][2] = c1;
      return true;
    case G_CZ:  // :169-174
      m[0][0] = c1; m[1][1] = c1; m[2][2] = c1; m[3][3] = -c1;
      return true;
    case G_SWAP:  // :177-182
      m[0][0] = c1; m[1][2] = c1; m[2][1] = c1; m[3][3] = c1;
      return true;
    case G_RZZ: {  // :185-195
      double half = g.theta / 2.0;
      C64 en = from_polar(1.0, -half), ep = from_polar(1.0, half);
      m[0][0] = en; m[1][1] = ep; m[2][2] = ep; m[3][3] = en;
      return true;
    }
    default:
      return false;
  }
}

// ---------------------------------------------------------------------------
// state.rs — QuantumState (f64)
// ---------------------------------------------------------------------------

// state.rs:214-229
void apply_1q(C64* a, size_t n, uint32_t qubit, const C64 m[2][2]) {
  size_t step = (size_t)1 << qubit;
  for (size_t block = 0; block < n; block += step << 1) {
    for (size_t i = block; i < block + step; ++i) {
      size_t j = i + step;
      C64 x = a[i], y = a[j];
      a[i] = m[0][0] * x + m[0][1] * y;
      a[j] = m[1][0] * x + m[1][1] * y;
    }
  }
}

// state.rs:238-276
void apply_2q(C64* a, size_t n, uint32_t q1, uint32_t q2, const C64 m[4][4]) {
  size_t b1 = (size_t)1 << q1, b2 = (size_t)1 << q2;
  for (size_t base = 0; base < n; ++base) {
    if ((base & b1) || (base & b2)) continue;
    size_t idx[4] = {base, base | b2, base | b1, base | b1 | b2};
    C64 v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
    for (int r = 0; r < 4; ++r)
      a[idx[r]] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2] + m[r][3] * v[3];
  }
}

// state.rs:288-326 — measure with a supplied uniform draw (state.rs:302).
void measure64(C64* a, size_t n, uint32_t qubit, double draw, int* result, double* prob_out) {
  size_t bit = (size_t)1 << qubit;
  double p0 = 0.0;
  for (size_t i = 0; i < n; ++i)
    if ((i & bit) == 0) p0 += norm_sq(a[i]);
  bool res = draw >= p0;
  double prob = res ? 1.0 - p0 : p0;
  double nf = prob > 0.0 ? 1.0 / std::sqrt(prob) : 0.0;
  for (size_t i = 0; i < n; ++i) {
    bool one = (i & bit) != 0;
    if (one == res) a[i] = scale(a[i], nf);
    else a[i] = C64{0.0, 0.0};
  }
  *result = res ? 1 : 0;
  *prob_out = prob;
}

// ---------------------------------------------------------------------------
// mixed_precision.rs — QuantumStateF32
// ---------------------------------------------------------------------------

// mixed_precision.rs:396-411
void apply_1q32(C32* a, size_t n, uint32_t qubit, const C32 m[2][2]) {
  size_t step = (size_t)1 << qubit;
  for (size_t block = 0; block < n; block += step << 1) {
    for (size_t i = block; i < block + step; ++i) {
      size_t j = i + step;
      C32 x = a[i], y = a[j];
      a[i] = m[0][0] * x + m[0][1] * y;
      a[j] = m[1][0] * x + m[1][1] * y;
    }
  }
}

// mixed_precision.rs:420-458
void apply_2q32(C32* a, size_t n, uint32_t q1, uint32_t q2, const C32 m[4][4]) {
  size_t b1 = (size_t)1 << q1, b2 = (size_t)1 << q2;
  for (size_t base = 0; base < n; ++base) {
    if ((base & b1) || (base & b2)) continue;
    size_t idx[4] = {base, base | b2, base | b1, base | b1 | b2};
    C32 v[4] = {a[idx[0]], a[idx[1]], a[idx[2]], a[idx[3]]};
    for (int r = 0; r < 4; ++r)
      a[idx[r]] = m[r][0] * v[0] + m[r][1] * v[1] + m[r][2] * v[2] + m[r][3] * v[3];
  }
}

// mixed_precision.rs:473-515 — p0 accumulated in f32, compared to an f64 draw.
void measure32(C32* a, size_t n, uint32_t qubit, double draw, int* result, double* prob_out) {
  size_t bit = (size_t)1 << qubit;
  float p0 = 0.0f;
  for (size_t i = 0; i < n; ++i)
    if ((i & bit) == 0) p0 += norm_sq32(a[i]);
  bool res = draw >= (double)p0;
  float prob = res ? 1.0f - p0 : p0;
  float nf = prob > 0.0f ? 1.0f / std::sqrt(prob) : 0.0f;
  for (size_t i = 0; i < n; ++i) {
    bool one = (i & bit) != 0;
    if (one == res) a[i] = scale32(a[i], nf);
    else a[i] = C32{0.0f, 0.0f};
  }
  *result = res ? 1 : 0;
  *prob_out = (double)prob;
}

void to32(const C64 m[2][2], C32 o[2][2]) {  // mixed_precision.rs:556-561
  for (int r = 0; r < 2; ++r)
    for (int c = 0; c < 2; ++c) o[r][c] = from_f64(m[r][c]);
}
void to32_4(const C64 m[4][4], C32 o[4][4]) {  // mixed_precision.rs:564-591
  for (int r = 0; r < 4; ++r)
    for (int c = 0; c < 4; ++c) o[r][c] = from_f64(m[r][c]);
}

int gate_qubits(const orc_gate& g, uint32_t q[2]) {  // gate.rs:40-64
  if (g.kind == G_BARRIER) return 0;
  if (is_2q_kind(g.kind)) { q[0] = g.q0; q[1] = g.q1; return 2; }
  q[0] = g.q0;
  return 1;
}

// state.rs:158-204 (f64) and mixed_precision.rs:336-386 (f32): apply_gate.
// `draws`/`ndraws`/`cursor` stand in for the state's StdRng.
int apply_gate_any(void* amps, int f32, uint32_t nq, const orc_gate& g, const double* draws,
                   size_t ndraws, size_t* cursor, orc_outcome* outs, size_t* nouts) {
  size_t n = (size_t)1 << nq;
  uint32_t qs[2];
  int k = gate_qubits(g, qs);
  for (int t = 0; t < k; ++t)  // state.rs:160-162, :441-449
    if (qs[t] >= nq) return E_INVALID_QUBIT;
  if (g.kind == G_BARRIER) return E_OK;  // :165
  if (g.kind == G_MEASURE || g.kind == G_RESET) {  // :167-175, :344-352
    if (*cursor >= ndraws) return E_CIRCUIT;
    double draw = draws[(*cursor)++];
    int res; double prob;
    if (f32) measure32((C32*)amps, n, g.q0, draw, &res, &prob);
    else measure64((C64*)amps, n, g.q0, draw, &res, &prob);
    if (g.kind == G_MEASURE) {
      if (outs) outs[*nouts] = orc_outcome{g.q0, res, prob};
      ++*nouts;
    } else if (res) {
      orc_gate x{}; x.kind = G_X; x.q0 = g.q0;
      C64 m[2][2]; matrix_1q(x, m);
      if (f32) { C32 m32[2][2]; to32(m, m32); apply_1q32((C32*)amps, n, g.q0, m32); }
      else apply_1q((C64*)amps, n, g.q0, m);
    }
    return E_OK;
  }
  if (is_2q_kind(g.kind)) {  // :178-188
    if (g.q0 == g.q1) return E_CIRCUIT;
    C64 m[4][4]; matrix_2q(g, m);
    if (f32) { C32 m32[4][4]; to32_4(m, m32); apply_2q32((C32*)amps, n, g.q0, g.q1, m32); }
    else apply_2q((C64*)amps, n, g.q0, g.q1, m);
    return E_OK;
  }
  C64 m[2][2];  // :191-202
  if (!matrix_1q(g, m)) return E_CIRCUIT;
  if (f32) { C32 m32[2][2]; to32(m, m32); apply_1q32((C32*)amps, n, g.q0, m32); }
  else apply_1q((C64*)amps, n, g.q0, m);
  return E_OK;
}

// state.rs:362-400 — <psi|P|psi> with ops in the given order.
double expectation64(const C64* a, size_t n, const uint32_t* qubits, const int32_t* ops, size_t nops) {
  C64 result{0.0, 0.0};
  for (size_t i = 0; i < n; ++i) {
    size_t j = i;
    C64 phase{1.0, 0.0};
    for (size_t t = 0; t < nops; ++t) {
      uint32_t q = qubits[t];
      size_t bit = (i >> q) & 1;
      switch (ops[t]) {
        case 0: break;  // I
        case 1: j ^= (size_t)1 << q; break;  // X
        case 2:  // Y
          j ^= (size_t)1 << q;
          phase = bit == 0 ? phase * C64{0.0, 1.0} : phase * C64{0.0, -1.0};
          break;
        case 3:  // Z
          if (bit == 1) phase = -phase;
          break;
      }
    }
    C64 t = conj(a[j]) * phase * a[i];
    result = C64{result.re + t.re, result.im + t.im};
  }
  return result.re;
}

// types.rs:207 PauliString terms, flattened: term k owns ops [off[k], off[k+1]).
double hamiltonian64(const C64* a, size_t n, const double* coeffs, const uint64_t* off,
                     const uint32_t* qubits, const int32_t* ops, size_t nterms) {
  double s = 0.0;  // state.rs:403-408: iterator sum, sequential
  for (size_t k = 0; k < nterms; ++k)
    s += coeffs[k] * expectation64(a, n, qubits + off[k], ops + off[k], off[k + 1] - off[k]);
  return s;
}

}  // namespace

// ===========================================================================
// C ABI for the tests
// ===========================================================================
extern "C" {

int orc_matrix_1q(const orc_gate* g, double out[8]) {
  C64 m[2][2];
  if (!matrix_1q(*g, m)) return -1;
  memcpy(out, m, sizeof(m));
  return 0;
}

int orc_matrix_2q(const orc_gate* g, double out[32]) {
  C64 m[4][4];
  if (!matrix_2q(*g, m)) return -1;
  memcpy(out, m, sizeof(m));
  return 0;
}

void orc_apply_1q_f64(double* amps, uint32_t nq, uint32_t q, const double m[8]) {
  C64 mm[2][2]; memcpy(mm, m, sizeof(mm));
  apply_1q((C64*)amps, (size_t)1 << nq, q, mm);
}

void orc_apply_2q_f64(double* amps, uint32_t nq, uint32_t q1, uint32_t q2, const double m[32]) {
  C64 mm[4][4]; memcpy(mm, m, sizeof(mm));
  apply_2q((C64*)amps, (size_t)1 << nq, q1, q2, mm);
}

void orc_apply_1q_f32(float* amps, uint32_t nq, uint32_t q, const float m[8]) {
  C32 mm[2][2]; memcpy(mm, m, sizeof(mm));
  apply_1q32((C32*)amps, (size_t)1 << nq, q, mm);
}

void orc_apply_2q_f32(float* amps, uint32_t nq, uint32_t q1, uint32_t q2, const float m[32]) {
  C32 mm[4][4]; memcpy(mm, m, sizeof(mm));
  apply_2q32((C32*)amps, (size_t)1 << nq, q1, q2, mm);
}

// state.rs:31-71 / mixed_precision.rs:206-255: |0..0>.  f32 != 0 selects
// QuantumStateF32 semantics.  Validation: new(0) errors, new_with_seed(0) is
// accepted for f64 (state.rs:55-71) but not for f32 (mixed_precision.rs:233-238).
int orc_state_check(uint32_t nq, int f32, int with_seed) {
  if (f32) {
    if (nq == 0) return E_CIRCUIT;
    if (nq > 33) return E_QUBIT_LIMIT;
    return E_OK;
  }
  if (nq == 0 && !with_seed) return E_CIRCUIT;
  if (nq > 32) return E_QUBIT_LIMIT;
  return E_OK;
}

int orc_apply_gate(void* amps, int f32, uint32_t nq, const orc_gate* g, const double* draws,
                   size_t ndraws, size_t* cursor, orc_outcome* outs, size_t* nouts) {
  return apply_gate_any(amps, f32, nq, *g, draws, ndraws, cursor, outs, nouts);
}

// simulator.rs:56-101 — run_with_config without noise: |0..0>, then every
// gate in order; gate_count counts unitary gates only (:73-75).
int orc_run_circuit(void* amps, int f32, uint32_t nq, const orc_gate* gates, size_t ngates,
                    const double* draws, size_t ndraws, orc_outcome* outs, size_t* nouts,
                    size_t* gate_count) {
  size_t n = (size_t)1 << nq;
  if (f32) {
    C32* a = (C32*)amps;
    for (size_t i = 0; i < n; ++i) a[i] = C32{0.0f, 0.0f};
    a[0] = C32{1.0f, 0.0f};
  } else {
    C64* a = (C64*)amps;
    for (size_t i = 0; i < n; ++i) a[i] = C64{0.0, 0.0};
    a[0] = C64{1.0, 0.0};
  }
  size_t cursor = 0, cnt = 0;
  *nouts = 0;
  for (size_t k = 0; k < ngates; ++k) {
    int rc = apply_gate_any(amps, f32, nq, gates[k], draws, ndraws, &cursor, outs, nouts);
    if (rc) return rc;
    if (!is_non_unitary(gates[k].kind)) ++cnt;
  }
  if (gate_count) *gate_count = cnt;
  return E_OK;
}

// simulator.rs:56-101 with a noise model: after every gate, apply_noise
// (simulator.rs:178-221) for each qubit the gate touches (Gate::qubits,
// gate.rs:40-64): depolarising (one draw; below the rate a second draw picks
// X / Y / Z at 1/3, 2/3), then bit flip (one draw, X), then phase flip (one
// draw, Z); a channel with rate 0 consumes nothing.  Measurement and noise
// draws come from one stream in that order.  *draws_used = draws consumed.
int orc_run_circuit_noisy(void* amps, int f32, uint32_t nq, const orc_gate* gates, size_t ngates,
                          double depol, double bit_flip, double phase_flip, const double* draws,
                          size_t ndraws, orc_outcome* outs, size_t* nouts, size_t* gate_count,
                          size_t* draws_used) {
  size_t n = (size_t)1 << nq;
  if (f32) {
    C32* a = (C32*)amps;
    for (size_t i = 0; i < n; ++i) a[i] = C32{0.0f, 0.0f};
    a[0] = C32{1.0f, 0.0f};
  } else {
    C64* a = (C64*)amps;
    for (size_t i = 0; i < n; ++i) a[i] = C64{0.0, 0.0};
    a[0] = C64{1.0, 0.0};
  }
  size_t cursor = 0, cnt = 0;
  *nouts = 0;
  auto pauli = [&](int kind, uint32_t q) {
    orc_gate p{};
    p.kind = kind;
    p.q0 = q;
    return apply_gate_any(amps, f32, nq, p, nullptr, 0, nullptr, nullptr, nullptr);
  };
  for (size_t k = 0; k < ngates; ++k) {
    int rc = apply_gate_any(amps, f32, nq, gates[k], draws, ndraws, &cursor, outs, nouts);
    if (rc) return rc;
    if (!is_non_unitary(gates[k].kind)) ++cnt;
    uint32_t qs[2];
    const int nqs = gate_qubits(gates[k], qs);
    for (int t = 0; t < nqs; ++t) {
      if (depol > 0.0) {
        if (cursor >= ndraws) return E_CIRCUIT;
        if (draws[cursor++] < depol) {
          if (cursor >= ndraws) return E_CIRCUIT;
          const double choice = draws[cursor++];
          rc = pauli(choice < 1.0 / 3.0 ? G_X : (choice < 2.0 / 3.0 ? G_Y : G_Z), qs[t]);
          if (rc) return rc;
        }
      }
      if (bit_flip > 0.0) {
        if (cursor >= ndraws) return E_CIRCUIT;
        if (draws[cursor++] < bit_flip && (rc = pauli(G_X, qs[t]))) return rc;
      }
      if (phase_flip > 0.0) {
        if (cursor >= ndraws) return E_CIRCUIT;
        if (draws[cursor++] < phase_flip && (rc = pauli(G_Z, qs[t]))) return rc;
      }
    }
  }
  if (gate_count) *gate_count = cnt;
  if (draws_used) *draws_used = cursor;
  return E_OK;
}

// grover.rs:69-74 optimal_iterations: max(round(pi/4 / theta - 0.5), 1) with
// theta = asin(sqrt(M / N)); Rust's f64::max ignores NaN (fmax), `as u32`
// saturates.
uint32_t orc_grover_optimal_iterations(uint32_t nq, size_t m) {
  const double n = (double)((size_t)1 << nq);
  const double theta = std::asin(std::sqrt((double)m / n));
  const double k = std::fmax(std::round(0.78539816339744830962 / theta - 0.5), 1.0);
  if (!(k < 4294967295.0)) return 4294967295u;
  return (uint32_t)k;
}

// grover.rs:98-183 run_grover (QuantumState, f64): H on every qubit; per
// iteration negate each listed target (:130-139), H on every qubit, negate
// every amplitude but index 0 (:146-155), H on every qubit; success
// probability = sum of probabilities()[t] over the list in order (:161-162);
// measure qubits 0..n-1 in turn with draws[q] (:187-200).  iterations < 0
// selects optimal_iterations.  amps_out receives the final (collapsed) state.
int orc_run_grover(uint32_t nq, const uint64_t* targets, size_t nt, int64_t iterations, const double* draws,
                   double* amps_out, uint64_t* measured, int* found, double* psucc, uint32_t* iters) {
  const size_t n = (size_t)1 << nq;
  for (size_t k = 0; k < nt; ++k)
    if (targets[k] >= n) return E_CIRCUIT;  // an assert (panic) in the reference
  const uint32_t it_count = iterations < 0 ? orc_grover_optimal_iterations(nq, nt) : (uint32_t)iterations;
  std::vector<C64> a(n, C64{0.0, 0.0});
  a[0] = C64{1.0, 0.0};
  orc_gate hg{};
  hg.kind = G_H;
  C64 h[2][2];
  matrix_1q(hg, h);
  auto hall = [&]() {
    for (uint32_t q = 0; q < nq; ++q) apply_1q(a.data(), n, q, h);
  };
  hall();
  for (uint32_t it = 0; it < it_count; ++it) {
    for (size_t k = 0; k < nt; ++k) a[targets[k]] = C64{-a[targets[k]].re, -a[targets[k]].im};
    hall();
    for (size_t i = 1; i < n; ++i) a[i] = C64{-a[i].re, -a[i].im};
    hall();
  }
  double ps = 0.0;
  for (size_t k = 0; k < nt; ++k) {
    const C64 x = a[targets[k]];
    ps += x.re * x.re + x.im * x.im;
  }
  uint64_t m = 0;
  for (uint32_t q = 0; q < nq; ++q) {
    int res;
    double p;
    measure64(a.data(), n, q, draws[q], &res, &p);
    if (res) m |= (uint64_t)1 << q;
  }
  *measured = m;
  *found = 0;
  for (size_t k = 0; k < nt; ++k)
    if (targets[k] == m) *found = 1;
  *psucc = ps;
  *iters = it_count;
  if (amps_out) memcpy(amps_out, a.data(), n * sizeof(C64));
  return E_OK;
}

// state.rs:288-326 / mixed_precision.rs:473-515
int orc_measure(void* amps, int f32, uint32_t nq, uint32_t q, double draw, int* result, double* prob) {
  if (q >= nq) return E_INVALID_QUBIT;
  if (f32) measure32((C32*)amps, (size_t)1 << nq, q, draw, result, prob);
  else measure64((C64*)amps, (size_t)1 << nq, q, draw, result, prob);
  return E_OK;
}

// state.rs:329-335 — measure_all, qubit 0 first, one draw per qubit.
int orc_measure_all(void* amps, int f32, uint32_t nq, const double* draws, int* results, double* probs) {
  for (uint32_t q = 0; q < nq; ++q) {
    int rc = orc_measure(amps, f32, nq, q, draws[q], &results[q], &probs[q]);
    if (rc) return rc;
  }
  return E_OK;
}

}  // extern "C"

namespace {

// measure_all (state.rs:329-335) of many shots of one state, memoised over the
// outcome prefixes the shots share.  After qubits 0..k-1 collapsed to the bits
// r, the reference's state is zero except at indices i = r + j 2^k, so `sub`
// keeps those amplitudes (j in order) and skips the zeros: adding norm_sq(0) =
// +0 to the non-negative running p0 leaves it unchanged, and a zero scaled by
// the norm factor stays zero.  Every other operation is measure64 /
// measure32's, in the same order, so each shot's bits equal those of
// measure_all on its own copy of the state.
template <class Amp, class Real>
void shots_level(std::vector<Amp>& sub, uint32_t k, uint32_t nq, const double* draws, const std::vector<uint32_t>& ids,
                 uint8_t* bits) {
  if (k == nq || ids.empty()) return;
  Real p0 = 0;
  for (size_t j = 0; j < sub.size(); j += 2) p0 += sub[j].re * sub[j].re + sub[j].im * sub[j].im;
  for (int b = 0; b < 2; ++b) {
    std::vector<uint32_t> grp;
    for (uint32_t s : ids)
      if ((draws[(size_t)s * nq + k] >= (double)p0) == (b == 1)) grp.push_back(s);
    if (grp.empty()) continue;
    for (uint32_t s : grp) bits[(size_t)s * nq + k] = (uint8_t)b;
    Real prob = b ? (Real)1 - p0 : p0;
    Real nf = prob > 0 ? (Real)1 / std::sqrt(prob) : (Real)0;
    std::vector<Amp> child(sub.size() / 2);
    for (size_t j = 0; j < child.size(); ++j) child[j] = Amp{sub[2 * j + b].re * nf, sub[2 * j + b].im * nf};
    shots_level<Amp, Real>(child, k + 1, nq, draws, grp, bits);
  }
}

}  // namespace

extern "C" {

// measure_all on a copy of `amps` for every shot (draws[shot*nq + q]), bits
// out as bits[shot*nq + q].  Same results as orc_measure_all per shot.
int orc_measure_all_shots(const void* amps, int f32, uint32_t nq, const double* draws, size_t shots, uint8_t* bits) {
  size_t n = (size_t)1 << nq;
  std::vector<uint32_t> ids(shots);
  for (size_t s = 0; s < shots; ++s) ids[s] = (uint32_t)s;
  if (f32) {
    std::vector<C32> sub((const C32*)amps, (const C32*)amps + n);
    shots_level<C32, float>(sub, 0, nq, draws, ids, bits);
  } else {
    std::vector<C64> sub((const C64*)amps, (const C64*)amps + n);
    shots_level<C64, double>(sub, 0, nq, draws, ids, bits);
  }
  return E_OK;
}

// simulator.rs:107-164 — run_shots, with draws supplied per shot
// (draws[shot*stride ...]).  Each shot re-simulates the circuit (:123-131);
// if the circuit has no Measure gate, measure_all follows (:133-137).
// bits[shot*nq + q] receives the last outcome recorded for qubit q (:139-145).
int orc_run_shots(int f32, uint32_t nq, const orc_gate* gates, size_t ngates, size_t shots,
                  const double* draws, size_t stride, uint8_t* bits) {
  size_t n = (size_t)1 << nq;
  size_t esz = f32 ? 8 : 16;
  std::vector<unsigned char> buf(n * esz);
  bool has_meas = false;
  for (size_t k = 0; k < ngates; ++k) has_meas |= gates[k].kind == G_MEASURE;
  std::vector<orc_outcome> outs(ngates + nq);
  for (size_t s = 0; s < shots; ++s) {
    size_t nouts = 0, cnt = 0;
    const double* d = draws + s * stride;
    // run_with_config consumes draws[0..m) for Measure/Reset gates
    size_t n_mr = 0;
    for (size_t k = 0; k < ngates; ++k) n_mr += gates[k].kind == G_MEASURE || gates[k].kind == G_RESET;
    int rc = orc_run_circuit(buf.data(), f32, nq, gates, ngates, d, n_mr, outs.data(), &nouts, &cnt);
    if (rc) return rc;
    if (!has_meas) {
      for (uint32_t q = 0; q < nq; ++q) {
        int r; double p;
        orc_measure(buf.data(), f32, nq, q, d[n_mr + q], &r, &p);
        outs[nouts++] = orc_outcome{q, r, p};
      }
    }
    uint8_t* b = bits + s * nq;
    memset(b, 0, nq);
    for (size_t t = 0; t < nouts; ++t)
      if (outs[t].qubit < nq) b[outs[t].qubit] = (uint8_t)outs[t].result;
  }
  return E_OK;
}

// state.rs:123-125
void orc_probabilities(const void* amps, int f32, uint32_t nq, double* out) {
  size_t n = (size_t)1 << nq;
  if (f32) {
    const C32* a = (const C32*)amps;
    for (size_t i = 0; i < n; ++i) out[i] = (double)norm_sq32(a[i]);  // mixed_precision.rs:303-305
  } else {
    const C64* a = (const C64*)amps;
    for (size_t i = 0; i < n; ++i) out[i] = norm_sq(a[i]);
  }
}

// state.rs:128-137
double orc_probability_of_qubit(const double* amps, uint32_t nq, uint32_t q) {
  const C64* a = (const C64*)amps;
  size_t n = (size_t)1 << nq, bit = (size_t)1 << q;
  double p1 = 0.0;
  for (size_t i = 0; i < n; ++i)
    if (i & bit) p1 += norm_sq(a[i]);
  return p1;
}

double orc_expectation(const double* amps, uint32_t nq, const uint32_t* qubits, const int32_t* ops, size_t nops) {
  return expectation64((const C64*)amps, (size_t)1 << nq, qubits, ops, nops);
}

double orc_expectation_hamiltonian(const double* amps, uint32_t nq, const double* coeffs, const uint64_t* off,
                                   const uint32_t* qubits, const int32_t* ops, size_t nterms) {
  return hamiltonian64((const C64*)amps, (size_t)1 << nq, coeffs, off, qubits, ops, nterms);
}

// state.rs:415-423
void orc_normalize(double* amps, uint32_t nq) {
  C64* a = (C64*)amps;
  size_t n = (size_t)1 << nq;
  double s = 0.0;
  for (size_t i = 0; i < n; ++i) s += norm_sq(a[i]);
  if (s > 0.0) {
    double inv = 1.0 / std::sqrt(s);
    for (size_t i = 0; i < n; ++i) a[i] = scale(a[i], inv);
  }
}

// state.rs:426-435
double orc_fidelity(const double* x, const double* y, uint32_t nq) {
  const C64* a = (const C64*)x;
  const C64* b = (const C64*)y;
  size_t n = (size_t)1 << nq;
  C64 inner{0.0, 0.0};
  for (size_t i = 0; i < n; ++i) {
    C64 t = conj(a[i]) * b[i];
    inner = C64{inner.re + t.re, inner.im + t.im};
  }
  return norm_sq(inner);
}

// circuit.rs:153-184 — depth().
uint32_t orc_circuit_depth(uint32_t nq, const orc_gate* gates, size_t ngates) {
  std::vector<uint32_t> d(nq, 0);
  for (size_t k = 0; k < ngates; ++k) {
    const orc_gate& g = gates[k];
    if (g.kind == G_BARRIER) {
      uint32_t mx = 0;
      for (uint32_t v : d) mx = v > mx ? v : mx;
      for (auto& v : d) v = mx;
      continue;
    }
    uint32_t qs[2];
    int c = gate_qubits(g, qs);
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

// optimizer.rs:12-23 — C = A*B.
void orc_mat_mul_2x2(const double a[8], const double b[8], double c[8]) {
  C64 A[2][2], B[2][2], Cm[2][2];
  memcpy(A, a, sizeof(A)); memcpy(B, b, sizeof(B));
  Cm[0][0] = A[0][0] * B[0][0] + A[0][1] * B[1][0];
  Cm[0][1] = A[0][0] * B[0][1] + A[0][1] * B[1][1];
  Cm[1][0] = A[1][0] * B[0][0] + A[1][1] * B[1][0];
  Cm[1][1] = A[1][0] * B[0][1] + A[1][1] * B[1][1];
  memcpy(c, Cm, sizeof(Cm));
}

// optimizer.rs:46-94 — fuse_gates: greedy fusion of adjacent 1q unitaries on
// the same qubit, fused = next * fused.  Returns the output gate count.
size_t orc_fuse_gates(const orc_gate* g, size_t len, orc_gate* out) {
  size_t i = 0, o = 0;
  while (i < len) {
    C64 first[2][2];
    if (!is_non_unitary(g[i].kind) && matrix_1q(g[i], first)) {
      uint32_t q = g[i].q0;
      C64 fused[2][2]; memcpy(fused, first, sizeof(fused));
      size_t count = 1;
      while (i + count < len) {
        const orc_gate& nx = g[i + count];
        if (is_non_unitary(nx.kind)) break;
        C64 nm[2][2];
        if (matrix_1q(nx, nm) && nx.q0 == q) {
          double c[8];
          orc_mat_mul_2x2((const double*)nm, (const double*)fused, c);
          memcpy(fused, c, sizeof(fused));
          ++count;
          continue;
        }
        break;
      }
      if (count > 1) {
        orc_gate u{}; u.kind = G_UNITARY1Q; u.q0 = q;
        memcpy(u.u, fused, sizeof(fused));
        out[o++] = u;
      } else {
        out[o++] = g[i];
      }
      i += count;
      continue;
    }
    out[o++] = g[i];
    ++i;
  }
  return o;
}

// ===========================================================================
// ruqu-algorithms: QAOA (qaoa.rs) and VQE (vqe.rs)
// ===========================================================================

// qaoa.rs:125-153 — build_qaoa_circuit; returns the gate count written.
size_t orc_build_qaoa_circuit(uint32_t nq, const uint32_t* ei, const uint32_t* ej, const double* w, size_t ne,
                              const double* gammas, const double* betas, uint32_t p, orc_gate* out) {
  size_t o = 0;
  for (uint32_t q = 0; q < nq; ++q) { orc_gate h{}; h.kind = G_H; h.q0 = q; out[o++] = h; }
  for (uint32_t l = 0; l < p; ++l) {
    for (size_t e = 0; e < ne; ++e) {
      orc_gate r{}; r.kind = G_RZZ; r.q0 = ei[e]; r.q1 = ej[e];
      r.theta = 2.0 * gammas[l] * w[e];
      out[o++] = r;
    }
    for (uint32_t q = 0; q < nq; ++q) {
      orc_gate r{}; r.kind = G_RX; r.q0 = q; r.theta = 2.0 * betas[l];
      out[o++] = r;
    }
  }
  return o;
}

// qaoa.rs:162-169
double orc_cut_value(const uint32_t* ei, const uint32_t* ej, const double* w, size_t ne, const uint8_t* bits) {
  double s = 0.0;
  for (size_t e = 0; e < ne; ++e)
    if (bits[ei[e]] != bits[ej[e]]) s += w[e];
  return s;
}

// qaoa.rs:179-201 — evaluate_qaoa_cost (f64 Simulator path).
// If amps_out is non-null it receives the final state.
int orc_qaoa_cost(uint32_t nq, const uint32_t* ei, const uint32_t* ej, const double* w, size_t ne,
                  const double* gammas, const double* betas, uint32_t p, double* cost, double* amps_out) {
  std::vector<orc_gate> gates(nq + (size_t)p * (ne + nq));
  size_t ng = orc_build_qaoa_circuit(nq, ei, ej, w, ne, gammas, betas, p, gates.data());
  size_t n = (size_t)1 << nq;
  std::vector<C64> st(n);
  size_t nouts = 0, cnt = 0;
  int rc = orc_run_circuit(st.data(), 0, nq, gates.data(), ng, nullptr, 0, nullptr, &nouts, &cnt);
  if (rc) return rc;
  double c = 0.0;
  for (size_t e = 0; e < ne; ++e) {
    uint32_t qs[2] = {ei[e], ej[e]};
    int32_t ops[2] = {3, 3};
    double zz = expectation64(st.data(), n, qs, ops, 2);
    c += w[e] * 0.5 * (1.0 - zz);
  }
  *cost = c;
  if (amps_out) memcpy(amps_out, st.data(), n * sizeof(C64));
  return E_OK;
}

typedef struct {
  double best_cut_value;
  uint32_t iterations;   // len(energy_history)
  int32_t converged;
  uint64_t best_index;   // index whose bits form best_bitstring
} orc_qaoa_result;

// qaoa.rs:217-308 — run_qaoa: gamma = beta = 0.5 start, pi/2 shifts, gammas
// updated before the beta gradients are taken, convergence at |dC| < 1e-6.
// gammas/betas are in/out (in: ignored, out: final angles); history has room
// for max_iter entries.
int orc_run_qaoa(uint32_t nq, const uint32_t* ei, const uint32_t* ej, const double* w, size_t ne,
                 uint32_t p, uint32_t max_iter, double lr, double* gammas, double* betas, double* history,
                 orc_qaoa_result* res) {
  std::vector<double> g(p, 0.5), b(p, 0.5);
  double best = -INFINITY;
  uint64_t best_idx = 0;
  bool conv = false;
  uint32_t hist = 0;
  size_t n = (size_t)1 << nq;
  std::vector<double> amps(2 * n);
  const double shift = 1.57079632679489661923132169163975144;  // FRAC_PI_2
  auto argmax_last = [&](const double* amps_in) {  // :246-253, Iterator::max_by keeps the LAST maximum
    const C64* a = (const C64*)amps_in;
    double bp = norm_sq(a[0]);
    uint64_t k = 0;
    for (size_t i = 1; i < n; ++i) {
      double pi = norm_sq(a[i]);
      if (ORC_PLANT == 4 ? (pi > bp) : !(pi < bp)) { bp = pi; k = i; }
    }
    return k;
  };
  for (uint32_t it = 0; it < max_iter; ++it) {
    double cost;
    int rc = orc_qaoa_cost(nq, ei, ej, w, ne, g.data(), b.data(), p, &cost, amps.data());
    if (rc) return rc;
    history[hist++] = cost;
    bool improved = cost > best;
    if (improved) {  // :238-259
      best = cost;
      best_idx = argmax_last(amps.data());
    }
    const double prev = ORC_PLANT == 5 ? history[0] : (it > 0 ? history[it - 1] : 0.0);
    if (it > 0 && std::fabs(cost - prev) < 1e-6) { conv = true; break; }  // :264-270
    const std::vector<double> g_old = g;
    for (uint32_t i = 0; i < p; ++i) {  // :278-286
      const std::vector<double>& base = ORC_PLANT == 2 ? g_old : g;
      std::vector<double> gp = base, gm = base;
      const uint32_t k = ORC_PLANT == 3 ? (i + 1) % p : i;
      gp[k] += shift; gm[k] -= shift;
      double cp, cm;
      orc_qaoa_cost(nq, ei, ej, w, ne, gp.data(), b.data(), p, &cp, nullptr);
      orc_qaoa_cost(nq, ei, ej, w, ne, gm.data(), b.data(), p, &cm, nullptr);
      if (ORC_PLANT == 1) g[i] -= lr * (cp - cm) / 2.0;
      else g[i] += lr * (cp - cm) / 2.0;
    }
    for (uint32_t i = 0; i < p; ++i) {  // :289-297
      std::vector<double> bp = b, bm = b;
      bp[i] += shift; bm[i] -= shift;
      double cp, cm;
      orc_qaoa_cost(nq, ei, ej, w, ne, g.data(), bp.data(), p, &cp, nullptr);
      orc_qaoa_cost(nq, ei, ej, w, ne, g.data(), bm.data(), p, &cm, nullptr);
      b[i] += lr * (cp - cm) / 2.0;
    }
    if (ORC_PLANT == 6 && improved) {
      double c2;
      orc_qaoa_cost(nq, ei, ej, w, ne, g.data(), b.data(), p, &c2, amps.data());
      best_idx = argmax_last(amps.data());
    }
  }
  memcpy(gammas, g.data(), p * sizeof(double));
  memcpy(betas, b.data(), p * sizeof(double));
  res->best_cut_value = best;
  res->iterations = hist;
  res->converged = conv ? 1 : 0;
  res->best_index = best_idx;
  return E_OK;
}

// vqe.rs:81-112 — build_ansatz (Ry layer, Rz layer, CNOT chain per layer).
size_t orc_build_ansatz(uint32_t nq, uint32_t depth, const double* params, orc_gate* out) {
  size_t o = 0, idx = 0;
  for (uint32_t l = 0; l < depth; ++l) {
    if (ORC_PLANT == 12) idx = 0;
    const int first = ORC_PLANT == 11 ? G_RZ : G_RY, second = ORC_PLANT == 11 ? G_RY : G_RZ;
    for (uint32_t q = 0; q < nq; ++q) { orc_gate g{}; g.kind = first; g.q0 = q; g.theta = params[idx++]; out[o++] = g; }
    for (uint32_t q = 0; q < nq; ++q) { orc_gate g{}; g.kind = second; g.q0 = q; g.theta = params[idx++]; out[o++] = g; }
    for (uint32_t q = 0; q + 1 < nq; ++q) {
      orc_gate g{}; g.kind = G_CNOT;
      g.q0 = ORC_PLANT == 10 ? q + 1 : q;
      g.q1 = ORC_PLANT == 10 ? q : q + 1;
      out[o++] = g;
    }
  }
  return o;
}

// vqe.rs:122-131 — evaluate_energy for an arbitrary gate list (the ansatz is
// built by the caller, so the Ry-CZ ansatz of the batched workload uses the
// same path).
int orc_energy(uint32_t nq, const orc_gate* gates, size_t ngates, const double* coeffs, const uint64_t* off,
               const uint32_t* qubits, const int32_t* ops, size_t nterms, double* energy) {
  size_t n = (size_t)1 << nq;
  std::vector<C64> st(n);
  size_t nouts = 0, cnt = 0;
  int rc = orc_run_circuit(st.data(), 0, nq, gates, ngates, nullptr, 0, nullptr, &nouts, &cnt);
  if (rc) return rc;
  *energy = hamiltonian64(st.data(), n, coeffs, off, qubits, ops, nterms);
  return E_OK;
}

typedef struct {
  double optimal_energy;
  uint32_t num_iterations;
  int32_t converged;
} orc_vqe_result;

// vqe.rs:148-216 — run_vqe with build_ansatz: params start at 0.1, simultaneous
// parameter-shift gradient, best tracked as the minimum.
int orc_run_vqe(uint32_t nq, uint32_t depth, const double* coeffs, const uint64_t* off, const uint32_t* qubits,
                const int32_t* ops, size_t nterms, uint32_t max_iter, double conv_thr, double lr,
                double* best_params, double* history, orc_vqe_result* res) {
  size_t np = (size_t)2 * nq * depth;  // vqe.rs:65-67
  std::vector<double> params(np, 0.1), bestp = params;
  double best = 1.7976931348623157e308;  // f64::MAX
  bool conv = false;
  uint32_t hist = 0;
  std::vector<orc_gate> gates((size_t)depth * (2 * nq + nq));
  const double shift = 1.57079632679489661923132169163975144;
  auto energy = [&](const std::vector<double>& pr, double* e) {
    size_t ng = orc_build_ansatz(nq, depth, pr.data(), gates.data());
    return orc_energy(nq, gates.data(), ng, coeffs, off, qubits, ops, nterms, e);
  };
  for (uint32_t it = 0; it < max_iter; ++it) {
    double e;
    int rc = energy(params, &e);
    if (rc) return rc;
    history[hist++] = e;
    if (ORC_PLANT == 9 ? (hist == 1 || e > best) : e < best) { best = e; bestp = params; }
    if (it > 0 && std::fabs(history[it - 1] - e) < conv_thr) { conv = true; break; }
    std::vector<double> grad(np, 0.0);
    for (size_t i = 0; i < np; ++i) {  // :188-198
      std::vector<double> pp = params, pm = params;
      pp[i] += shift; pm[i] -= shift;
      double ep, em;
      energy(pp, &ep);
      energy(pm, &em);
      grad[i] = (ep - em) / 2.0;
      if (ORC_PLANT == 8) params[i] -= lr * grad[i];
    }
    if (ORC_PLANT != 8)  // :203-205, after every gradient entry is known
      for (size_t i = 0; i < np; ++i) params[i] += (ORC_PLANT == 7 ? lr : -lr) * grad[i];
  }
  memcpy(best_params, bestp.data(), np * sizeof(double));
  res->optimal_energy = best;
  res->num_iterations = hist;
  res->converged = conv ? 1 : 0;
  return E_OK;
}

}  // extern "C"
