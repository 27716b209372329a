// Warm-start graph network for QAOA angles (no reference counterpart): sizes
// and parameter layout shared by the host driver (mpnn.cpp) and the kernels
// (kernels_mpnn.hip).  The same layout is restated by oracle/mpnn_oracle.py.
#pragma once

#include <cstddef>
#include <cstdint>

namespace rq {
namespace mpnn {

constexpr int H = 256;        // hidden width
constexpr int KIN = 32;       // node feature columns (2 used, zero padded)
constexpr int GF = 32;        // graph feature columns (onehot(p), n/22, density)
constexpr int HIN = H + GF;   // head input
constexpr int NOUT = 64;      // head output columns (gammas 0..7, betas 8..15)
constexpr int PMAX = 8;
constexpr int LAYERS = 3;

// Flat FP32 parameter layout: W_in, b_in, (W_k, b_k) k=1..3, W_h1, b_h1, W_h2, b_h2.
// Weight matrices are row-major [rows = input features][cols = outputs].
struct Tensor {
  size_t off;
  int rows, cols;  // cols = 1 row-vector bias: rows = length
};

constexpr size_t OFF_W_IN = 0;
constexpr size_t OFF_B_IN = OFF_W_IN + (size_t)KIN * H;
constexpr size_t OFF_W1 = OFF_B_IN + H;
constexpr size_t LAYER_STRIDE = (size_t)2 * H * H + H;
constexpr size_t OFF_W_H1 = OFF_W1 + LAYERS * LAYER_STRIDE;
constexpr size_t OFF_B_H1 = OFF_W_H1 + (size_t)HIN * H;
constexpr size_t OFF_W_H2 = OFF_B_H1 + H;
constexpr size_t OFF_B_H2 = OFF_W_H2 + (size_t)H * NOUT;
constexpr size_t NPARAMS = OFF_B_H2 + NOUT;

constexpr size_t off_w(int k) { return OFF_W1 + (size_t)k * LAYER_STRIDE; }
constexpr size_t off_b(int k) { return off_w(k) + (size_t)2 * H * H; }

}  // namespace mpnn
}  // namespace rq
