// Dense-block execution of a circuit (rq_run_circuit with fusion = 2):
// gates are folded greedily, in circuit order, into blocks of at most 5
// qubits; each block's 32x32 complex unitary is built on the host in FP64
// from gate.rs's matrices (a block with fewer qubits is padded with unused
// qubits, on which it is the identity) and applied by one pass of the
// matrix-core kernel in kernels_dense.hip.  Complex-f32 states of at least
// 10 qubits only; other states fall back to the fused tile schedule.
// Measure/Reset gates split the circuit and run exactly, as in run_fused.
#include <algorithm>
#include <complex>
#include <cstring>
#include <vector>

#include "fusion.h"
#include "gates.h"

namespace rq {

void launch_dense5_f32(State& s, const int* q, const float* d_wt);

namespace {

using cd = std::complex<double>;

struct DenseBlock {
  int q[5];
  std::vector<cd> U;  // 32 x 32 row-major, index bit i <-> q[i]
};

struct Builder {
  std::vector<int> pos;  // qubit of each bit position in use
  std::vector<cd> U = std::vector<cd>(32 * 32);
  Builder() { reset(); }
  void reset() {
    pos.clear();
    std::fill(U.begin(), U.end(), cd(0.0, 0.0));
    for (int i = 0; i < 32; ++i) U[i * 32 + i] = 1.0;
  }
  int bit_of(int q) const {
    for (size_t i = 0; i < pos.size(); ++i)
      if (pos[i] == q) return (int)i;
    return -1;
  }
  int needed(const rq_gate& g) const {
    int k = 0;
    if (bit_of((int)g.q0) < 0) ++k;
    if (is_2q(g.kind) && bit_of((int)g.q1) < 0) ++k;
    return k;
  }
  int take(int q) {
    int b = bit_of(q);
    if (b >= 0) return b;
    pos.push_back(q);
    return (int)pos.size() - 1;
  }
  // U <- G U with G acting on the gate's bit positions.
  void apply(const rq_gate& g) {
    if (is_1q(g.kind)) {
      rq_c64 m[4];
      matrix_1q(g, m);
      const int j = take((int)g.q0);
      const cd a(m[0].re, m[0].im), b(m[1].re, m[1].im), c(m[2].re, m[2].im), d(m[3].re, m[3].im);
      for (int r0 = 0; r0 < 32; ++r0) {
        if (r0 & (1 << j)) continue;
        const int r1 = r0 | (1 << j);
        for (int col = 0; col < 32; ++col) {
          const cd u0 = U[r0 * 32 + col], u1 = U[r1 * 32 + col];
          U[r0 * 32 + col] = a * u0 + b * u1;
          U[r1 * 32 + col] = c * u0 + d * u1;
        }
      }
    } else {
      rq_c64 m[16];
      matrix_2q(g, m);
      const int j1 = take((int)g.q0), j2 = take((int)g.q1);
      for (int r = 0; r < 32; ++r) {
        if (r & ((1 << j1) | (1 << j2))) continue;
        int rows[4];
        for (int t = 0; t < 4; ++t) rows[t] = r | (((t >> 1) & 1) << j1) | ((t & 1) << j2);  // t = q1_bit*2 + q2_bit
        for (int col = 0; col < 32; ++col) {
          cd v[4], w[4];
          for (int t = 0; t < 4; ++t) v[t] = U[rows[t] * 32 + col];
          for (int t = 0; t < 4; ++t) {
            w[t] = 0.0;
            for (int u = 0; u < 4; ++u) w[t] += cd(m[t * 4 + u].re, m[t * 4 + u].im) * v[u];
          }
          for (int t = 0; t < 4; ++t) U[rows[t] * 32 + col] = w[t];
        }
      }
    }
  }
  // Pads to 5 qubits (identity on the added ones) and sorts the qubits
  // ascending, permuting the index bits of U accordingly.
  DenseBlock finish(uint32_t nq) const {
    std::vector<int> qs = pos;
    for (int q = 0; (int)qs.size() < 5 && q < (int)nq; ++q)
      if (std::find(qs.begin(), qs.end(), q) == qs.end()) qs.push_back(q);
    std::vector<int> order(5);
    for (int i = 0; i < 5; ++i) order[i] = i;
    std::sort(order.begin(), order.end(), [&](int x, int y) { return qs[x] < qs[y]; });
    // new bit i holds old bit order[i]
    auto remap = [&](int idx_new) {
      int old = 0;
      for (int i = 0; i < 5; ++i) old |= ((idx_new >> i) & 1) << order[i];
      return old;
    };
    DenseBlock blk;
    for (int i = 0; i < 5; ++i) blk.q[i] = qs[order[i]];
    blk.U.resize(32 * 32);
    for (int r = 0; r < 32; ++r)
      for (int c = 0; c < 32; ++c) blk.U[r * 32 + c] = U[remap(r) * 32 + remap(c)];
    return blk;
  }
};

void plan_dense(uint32_t nq, const rq_gate* g, size_t ng, std::vector<DenseBlock>& out, uint64_t* unitary) {
  Builder b;
  bool open = false;
  for (size_t k = 0; k < ng; ++k) {
    if (is_non_unitary(g[k].kind)) continue;  // barriers (measure/reset are split off by the caller)
    ++*unitary;
    if (open && (int)b.pos.size() + b.needed(g[k]) > 5) {
      out.push_back(b.finish(nq));
      b.reset();
    }
    b.apply(g[k]);
    open = true;
  }
  if (open) out.push_back(b.finish(nq));
}

// B operands of the matrix-core kernel: wt[(h*32 + s)*64 + l] = W[32h + l%32][2s + l/32]
// with W = [[Re U, -Im U], [Im U, Re U]].
void embed(const DenseBlock& blk, float* wt) {
  auto W = [&](int n, int k) -> double {
    const int r = n & 31, c = k & 31;
    const cd u = blk.U[r * 32 + c];
    if (n < 32) return k < 32 ? u.real() : -u.imag();
    return k < 32 ? u.imag() : u.real();
  };
  for (int h = 0; h < 2; ++h)
    for (int s = 0; s < 32; ++s)
      for (int l = 0; l < 64; ++l) wt[(h * 32 + s) * 64 + l] = (float)W(32 * h + (l & 31), 2 * s + (l >> 5));
}

int run_segment(State& s, const rq_gate* g, size_t ng, FusionStats* st) {
  std::vector<DenseBlock> blocks;
  uint64_t unitary = 0;
  plan_dense(s.nq, g, ng, blocks, &unitary);
  st->unitary += unitary;
  if (blocks.empty()) return RQ_OK;
  const size_t per = 64 * 64;
  std::vector<float> host(per * blocks.size());
  for (size_t i = 0; i < blocks.size(); ++i) embed(blocks[i], host.data() + i * per);
  const size_t bytes = host.size() * sizeof(float);
  invalidate_fused_program(s);  // s.prog is about to hold the dense blocks
  if (s.prog_cap < bytes) {
    if (s.prog) RQ_HIP(hipFree(s.prog));
    s.prog = nullptr;
    RQ_HIP(hipMalloc(&s.prog, bytes));
    s.prog_cap = bytes;
  }
  RQ_HIP(hipMemcpyAsync(s.prog, host.data(), bytes, hipMemcpyHostToDevice, s.stream));
  for (size_t i = 0; i < blocks.size(); ++i) launch_dense5_f32(s, blocks[i].q, (const float*)s.prog + i * per);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipStreamSynchronize(s.stream));  // the host staging buffer goes out of scope
  st->passes += blocks.size();
  st->hbm_bytes += 2.0 * (double)s.n * (double)s.amp_bytes() * (double)blocks.size();
  s.gate_count += unitary;
  return RQ_OK;
}

}  // namespace

int run_dense(State& s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws, size_t* cursor,
              rq_measurement* meas, size_t cap, size_t* produced, FusionStats* st) {
  if (s.prec != RQ_C32 || s.nq < 10) return run_fused(s, gates, ngates, draws, ndraws, cursor, meas, cap, produced, st);
  for (size_t k = 0; k < ngates; ++k)
    if (!is_non_unitary(gates[k].kind)) {
      const rq_gate& g = gates[k];
      if (g.q0 >= s.nq || (is_2q(g.kind) && (g.q1 >= s.nq || g.q1 == g.q0)))
        return run_fused(s, gates, ngates, draws, ndraws, cursor, meas, cap, produced, st);  // its checks report
    }
  size_t k = 0;
  while (k < ngates) {
    size_t e = k;
    while (e < ngates && gates[e].kind != RQ_GATE_MEASURE && gates[e].kind != RQ_GATE_RESET) ++e;
    if (e > k) {
      int rc = run_segment(s, gates + k, e - k, st);
      if (rc) return rc;
    }
    if (e < ngates) {
      const rq_gate& g = gates[e];
      if (*cursor >= ndraws) return fail(RQ_ERR_CIRCUIT, "measurement needs a uniform draw but the draw list is exhausted");
      const double draw = draws[(*cursor)++];
      if (g.q0 >= s.nq)
        return fail(RQ_ERR_INVALID_QUBIT, "invalid qubit index " + std::to_string(g.q0) + " for " +
                                              std::to_string(s.nq) + "-qubit system");
      int rc;
      if (g.kind == RQ_GATE_MEASURE) {
        rq_measurement mm;
        rc = measure_state(s, g.q0, draw, &mm);
        if (!rc) {
          if (meas && *produced < cap) meas[*produced] = mm;
          ++*produced;
        }
      } else {
        rc = reset_state(s, g.q0, draw);
      }
      if (rc) return rc;
    }
    k = e + 1;
  }
  st->fused = st->unitary > st->passes ? st->unitary - st->passes : 0;
  return RQ_OK;
}

}  // namespace rq
