// gate.rs closed forms, evaluated on the host (see gates.h).
#include "gates.h"

#include <math.h>

#include <cmath>

namespace rq {

namespace {
constexpr double FRAC_1_SQRT_2 = 0.70710678118654752440084436210484903928483593768847;
// cos/sin pairs of one argument go through one glibc sincos call: the
// reference's f64::cos/sin pair lowers to an LLVM `sincos` libcall on GNU
// targets, and glibc's sincos can differ from separate sin/cos in the last bit.
inline void cos_sin(double th, double* c, double* s) { ::sincos(th, s, c); }
inline rq_c64 polar1(double th) {  // types.rs:24-29 from_polar(1.0, th)
  double c, s;
  cos_sin(th, &c, &s);
  return rq_c64{1.0 * c, 1.0 * s};
}
}  // namespace

bool matrix_1q(const rq_gate& g, rq_c64 m[4]) {
  const rq_c64 c0{0.0, 0.0}, c1{1.0, 0.0}, ci{0.0, 1.0}, nci{-0.0, -1.0}, nc1{-1.0, -0.0};
  switch (g.kind) {
    case RQ_GATE_H: {
      rq_c64 h{FRAC_1_SQRT_2, 0.0}, nh{-FRAC_1_SQRT_2, -0.0};
      m[0] = h; m[1] = h; m[2] = h; m[3] = nh;
      return true;
    }
    case RQ_GATE_X: m[0] = c0; m[1] = c1; m[2] = c1; m[3] = c0; return true;
    case RQ_GATE_Y: m[0] = c0; m[1] = nci; m[2] = ci; m[3] = c0; return true;
    case RQ_GATE_Z: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = nc1; return true;
    case RQ_GATE_S: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = ci; return true;
    case RQ_GATE_SDG: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = nci; return true;
    case RQ_GATE_T: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = rq_c64{FRAC_1_SQRT_2, FRAC_1_SQRT_2}; return true;
    case RQ_GATE_TDG: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = rq_c64{FRAC_1_SQRT_2, -FRAC_1_SQRT_2}; return true;
    case RQ_GATE_RX: {
      double half = g.theta / 2.0, ch, sh;
      cos_sin(half, &ch, &sh);
      rq_c64 c{ch, 0.0}, s{0.0, -sh};
      m[0] = c; m[1] = s; m[2] = s; m[3] = c;
      return true;
    }
    case RQ_GATE_RY: {
      double half = g.theta / 2.0, ch, sh;
      cos_sin(half, &ch, &sh);
      m[0] = rq_c64{ch, 0.0}; m[1] = rq_c64{-sh, 0.0}; m[2] = rq_c64{sh, 0.0}; m[3] = rq_c64{ch, 0.0};
      return true;
    }
    case RQ_GATE_RZ: {
      double half = g.theta / 2.0;
      m[0] = polar1(-half); m[1] = c0; m[2] = c0; m[3] = polar1(half);
      return true;
    }
    case RQ_GATE_PHASE: m[0] = c1; m[1] = c0; m[2] = c0; m[3] = polar1(g.theta); return true;
    case RQ_GATE_UNITARY1Q:
      m[0] = g.u[0][0]; m[1] = g.u[0][1]; m[2] = g.u[1][0]; m[3] = g.u[1][1];
      return true;
    default:
      return false;
  }
}

bool matrix_2q(const rq_gate& g, rq_c64 m[16]) {
  const rq_c64 c0{0.0, 0.0}, c1{1.0, 0.0};
  for (int k = 0; k < 16; ++k) m[k] = c0;
  switch (g.kind) {
    case RQ_GATE_CNOT: m[0] = c1; m[5] = c1; m[11] = c1; m[14] = c1; return true;
    case RQ_GATE_CZ: m[0] = c1; m[5] = c1; m[10] = c1; m[15] = rq_c64{-1.0, -0.0}; return true;
    case RQ_GATE_SWAP: m[0] = c1; m[6] = c1; m[9] = c1; m[15] = c1; return true;
    case RQ_GATE_RZZ: {
      double half = g.theta / 2.0;
      rq_c64 en = polar1(-half), ep = polar1(half);
      m[0] = en; m[5] = ep; m[10] = ep; m[15] = en;
      return true;
    }
    default:
      return false;
  }
}

}  // namespace rq
