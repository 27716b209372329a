// Host-side gate algebra: the closed-form 2x2 / 4x4 matrices of gate.rs and
// helpers for classifying gates.  Matrices are evaluated on the host with the
// C library's cos/sin (the values Rust's f64::cos/sin return on the same
// libm) and uploaded; device code never evaluates a transcendental for a gate.
#pragma once

#include <cstdint>

#include "../../include/ruqu_gpu.h"

namespace rq {

inline bool is_2q(int k) { return k >= RQ_GATE_CNOT && k <= RQ_GATE_RZZ; }
inline bool is_1q(int k) { return (k >= RQ_GATE_H && k <= RQ_GATE_PHASE) || k == RQ_GATE_UNITARY1Q; }
inline bool is_non_unitary(int k) {  // gate.rs:67-69
  return k == RQ_GATE_MEASURE || k == RQ_GATE_RESET || k == RQ_GATE_BARRIER;
}
inline bool is_diagonal_1q(int k) {
  return k == RQ_GATE_Z || k == RQ_GATE_S || k == RQ_GATE_SDG || k == RQ_GATE_T || k == RQ_GATE_TDG ||
         k == RQ_GATE_RZ || k == RQ_GATE_PHASE;
}

// gate.rs:72-148.  m is row-major [2][2].  Returns false for non-1q gates.
bool matrix_1q(const rq_gate& g, rq_c64 m[4]);
// gate.rs:154-199.  m is row-major [4][4], index q1_bit*2 + q2_bit.
bool matrix_2q(const rq_gate& g, rq_c64 m[16]);

// Complex helpers with the reference's operation order (types.rs:81-90).
inline rq_c64 cmul(rq_c64 a, rq_c64 b) { return rq_c64{a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; }
inline rq_c64 cadd(rq_c64 a, rq_c64 b) { return rq_c64{a.re + b.re, a.im + b.im}; }

}  // namespace rq
