// Single-instance QAOA MaxCut driver on one state (qaoa.rs restated on top of
// the device kernels).
//
// evaluate_qaoa_cost (qaoa.rs:179-201): |+>^n, then per layer one fused
// cost-phase pass (bit-identical to the Rzz sequence) and the Rx(2*beta)
// mixer, then the expected cut.
//   mode 0 (exact): the mixer is one exact Rx kernel per qubit (the
//     reference's arithmetic); the cost is sum_e w*0.5*(1 - <Z_i Z_j>) over
//     per-edge FP64 reductions, combined in edge order as qaoa.rs:193-199.
//     f64 amplitudes are bit-identical to the reference's.
//   mode 1 (fused): the mixer runs through the fused tile schedule (all Rx of a
//     layer in ~n/m HBM passes) and the cost is one FP64 pass of
//     sum_i |a_i|^2 C(i).  Rounding differs at the 1e-15 level.
// run_qaoa (qaoa.rs:217-308): the plain optimizer, step for step.
#include <chrono>
#include <cmath>
#include <vector>

#include "fusion.h"
#include "gates.h"
#include "qaoa.h"

namespace rq {

struct Qaoa {
  uint32_t n = 0;
  int prec = RQ_C64;
  std::vector<uint32_t> ei, ej;
  std::vector<double> w;
  State* st = nullptr;
  void* d_edges = nullptr;   // p layers of edge-factor records
  size_t d_edges_cap = 0;
  void* d_wedges = nullptr;  // (i, j, w) records for the energy
  uint64_t evaluations = 0;
  double device_ms = 0;
};

namespace {

constexpr double FRAC_1_SQRT_2 = 0.70710678118654752440084436210484903928483593768847;

// Amplitude of |+>^n built by n H gates on |0..0>: each H multiplies by h, in
// the state's precision (QuantumStateF32 narrows h to f32 first,
// mixed_precision.rs:556-561).
double plus_amplitude(uint32_t n, int prec) {
  if (prec == RQ_C32) {
    const float h = (float)FRAC_1_SQRT_2;
    float c = 1.0f;
    for (uint32_t q = 0; q < n; ++q) c = h * c;
    return (double)c;
  }
  double c = 1.0;
  for (uint32_t q = 0; q < n; ++q) c = FRAC_1_SQRT_2 * c;
  return c;
}

int upload_layers(Qaoa& q, const double* gammas, uint32_t p) {
  const int ne = (int)q.ei.size();
  const size_t rec = qaoa_edge_record_bytes(q.prec);
  const size_t bytes = (size_t)p * ne * rec;
  if (bytes > q.d_edges_cap) {
    if (q.d_edges) {
      RQ_HIP(hipStreamSynchronize(q.st->stream));
      RQ_HIP(hipFree(q.d_edges));
      q.d_edges = nullptr;
    }
    RQ_HIP(hipMalloc(&q.d_edges, bytes ? bytes : 1));
    q.d_edges_cap = bytes;
  }
  std::vector<unsigned char> host(bytes);
  std::vector<rq_c64> en(ne), ep(ne);
  for (uint32_t l = 0; l < p; ++l) {
    for (int e = 0; e < ne; ++e) {
      rq_gate g{};
      g.kind = RQ_GATE_RZZ;
      g.q0 = q.ei[e];
      g.q1 = q.ej[e];
      g.theta = 2.0 * gammas[l] * q.w[e];  // qaoa.rs:144
      rq_c64 m[16];
      matrix_2q(g, m);
      en[e] = m[0];
      ep[e] = m[5];
    }
    qaoa_pack_edges(q.prec, q.ei.data(), q.ej.data(), en.data(), ep.data(), ne, host.data() + (size_t)l * ne * rec);
  }
  if (bytes) {
    RQ_HIP(hipMemcpyAsync(q.d_edges, host.data(), bytes, hipMemcpyHostToDevice, q.st->stream));
    RQ_HIP(hipStreamSynchronize(q.st->stream));
  }
  return RQ_OK;
}

int evaluate(Qaoa& q, const double* gammas, const double* betas, uint32_t p, int mode, double* cost) {
  State& s = *q.st;
  const int ne = (int)q.ei.size();
  hipEvent_t e0, e1;
  RQ_HIP(hipEventCreate(&e0));
  RQ_HIP(hipEventCreate(&e1));
  int rc = upload_layers(q, gammas, p);
  if (rc) return rc;
  RQ_HIP(hipEventRecord(e0, s.stream));
  rc = qaoa_fill_uniform(s, plus_amplitude(q.n, q.prec));
  if (rc) return rc;
  const size_t rec = qaoa_edge_record_bytes(q.prec);
  std::vector<rq_gate> mixer(q.n);
  for (uint32_t l = 0; l < p && !rc; ++l) {
    if (ne) rc = qaoa_cost_phase(s, (const unsigned char*)q.d_edges + (size_t)l * ne * rec, ne);
    if (rc) break;
    for (uint32_t k = 0; k < q.n; ++k) {
      mixer[k] = rq_gate{};
      mixer[k].kind = RQ_GATE_RX;
      mixer[k].q0 = k;
      mixer[k].theta = 2.0 * betas[l];  // qaoa.rs:148
    }
    if (mode == 0 || q.n < (uint32_t)FUSE_R_BITS) {
      for (uint32_t k = 0; k < q.n; ++k) {
        rq_c64 m[4];
        matrix_1q(mixer[k], m);
        launch_apply_1q(s, k, m);
      }
      RQ_HIP(hipGetLastError());
    } else {
      FusedPlan plan;
      rc = plan_fused(q.n, mixer.data(), mixer.size(), fusion_config(), &plan);
      if (!rc) rc = launch_fused_plan(s, plan);
    }
  }
  if (rc) return rc;
  if (mode == 0) {
    std::vector<uint64_t> masks(ne);
    for (int e = 0; e < ne; ++e) masks[e] = (1ull << q.ei[e]) | (1ull << q.ej[e]);
    std::vector<double> zz(ne);
    if (ne) rc = reduce_zmasks(s, masks.data(), ne, zz.data());
    if (rc) return rc;
    double c = 0.0;
    for (int e = 0; e < ne; ++e) c += q.w[e] * 0.5 * (1.0 - zz[e]);  // qaoa.rs:193-199
    *cost = c;
  } else {
    rc = qaoa_energy(s, q.d_wedges, ne, cost);
    if (rc) return rc;
  }
  RQ_HIP(hipEventRecord(e1, s.stream));
  RQ_HIP(hipEventSynchronize(e1));
  float ms = 0;
  RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  q.device_ms += ms;
  ++q.evaluations;
  return RQ_OK;
}

}  // namespace
}  // namespace rq

using namespace rq;

extern "C" {

int rq_qaoa_create(const rq_graph* g, int precision, rq_qaoa** out) {
  if (!g || !out || (g->num_edges && (!g->ei || !g->ej || !g->w))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (g->num_edges > (uint32_t)QAOA_MAX_EDGES) return fail(RQ_ERR_ARGUMENT, "too many edges for one QAOA instance");
  for (uint32_t e = 0; e < g->num_edges; ++e) {  // qaoa.rs:56-60 (the reference panics)
    uint32_t bad = g->ei[e] >= g->num_nodes ? g->ei[e] : (g->ej[e] >= g->num_nodes ? g->ej[e] : UINT32_MAX);
    if (bad != UINT32_MAX) return fail(RQ_ERR_INVALID_QUBIT, "node index " + std::to_string(bad) + " out of range");
  }
  rq_state* hs = nullptr;
  int rc = rq_state_create(g->num_nodes, precision, 1, &hs);
  if (rc) return rc;
  Qaoa* q = new Qaoa();
  q->n = g->num_nodes;
  q->prec = precision;
  q->ei.assign(g->ei, g->ei + g->num_edges);
  q->ej.assign(g->ej, g->ej + g->num_edges);
  q->w.assign(g->w, g->w + g->num_edges);
  q->st = (State*)hs;
  size_t wb = qaoa_weight_record_bytes() * (g->num_edges ? g->num_edges : 1);
  std::vector<unsigned char> host(wb);
  qaoa_pack_weights(g->ei, g->ej, g->w, (int)g->num_edges, host.data());
  hipError_t e = hipMalloc(&q->d_wedges, wb);
  if (e == hipSuccess) e = hipMemcpy(q->d_wedges, host.data(), wb, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    rq_qaoa_destroy((rq_qaoa*)q);
    return hip_fail(e, "qaoa setup");
  }
  *out = (rq_qaoa*)q;
  return RQ_OK;
}

void rq_qaoa_destroy(rq_qaoa* h) {
  Qaoa* q = (Qaoa*)h;
  if (!q) return;
  if (q->st) (void)hipStreamSynchronize(q->st->stream);
  if (q->d_edges) (void)hipFree(q->d_edges);
  if (q->d_wedges) (void)hipFree(q->d_wedges);
  rq_state_destroy((rq_state*)q->st);
  delete q;
}

rq_state* rq_qaoa_state(rq_qaoa* h) { return h ? (rq_state*)((Qaoa*)h)->st : nullptr; }

int rq_qaoa_cost(rq_qaoa* h, const double* gammas, const double* betas, uint32_t p, int mode, double* cost) {
  Qaoa* q = (Qaoa*)h;
  if (!q || !cost || (p && (!gammas || !betas))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  return evaluate(*q, gammas, betas, p, mode, cost);
}

// qaoa.rs:217-308.  init_gammas/init_betas may be null (then 0.5 each, as
// the reference).  history needs max_iter entries.
int rq_qaoa_run(rq_qaoa* h, uint32_t p, uint32_t max_iter, double lr, int mode, const double* init_gammas,
                const double* init_betas, double* gammas_out, double* betas_out, double* history,
                rq_qaoa_result* res) {
  Qaoa* q = (Qaoa*)h;
  if (!q || !res || (max_iter && !history)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  auto t0 = std::chrono::steady_clock::now();
  q->evaluations = 0;
  q->device_ms = 0;
  std::vector<double> g(p, 0.5), b(p, 0.5);
  if (init_gammas) g.assign(init_gammas, init_gammas + p);
  if (init_betas) b.assign(init_betas, init_betas + p);
  double best = -INFINITY;
  uint64_t best_idx = 0;
  bool conv = false;
  uint32_t hist = 0;
  const double shift = 1.57079632679489661923132169163975144;  // FRAC_PI_2
  int rc = RQ_OK;
  for (uint32_t it = 0; it < max_iter; ++it) {
    double cost;
    rc = evaluate(*q, g.data(), b.data(), p, mode, &cost);
    if (rc) return rc;
    history[hist++] = cost;
    if (cost > best) {  // qaoa.rs:238-259: argmax of the probabilities of this state
      best = cost;
      rc = argmax_probability(*q->st, &best_idx, nullptr);
      if (rc) return rc;
    }
    if (it > 0 && std::fabs(cost - history[it - 1]) < 1e-6) {
      conv = true;
      break;
    }
    for (uint32_t i = 0; i < p; ++i) {
      std::vector<double> gp = g, gm = g;
      gp[i] += shift;
      gm[i] -= shift;
      double cp, cm;
      rc = evaluate(*q, gp.data(), b.data(), p, mode, &cp);
      if (!rc) rc = evaluate(*q, gm.data(), b.data(), p, mode, &cm);
      if (rc) return rc;
      g[i] += lr * (cp - cm) / 2.0;
    }
    for (uint32_t i = 0; i < p; ++i) {
      std::vector<double> bp = b, bm = b;
      bp[i] += shift;
      bm[i] -= shift;
      double cp, cm;
      rc = evaluate(*q, g.data(), bp.data(), p, mode, &cp);
      if (!rc) rc = evaluate(*q, g.data(), bm.data(), p, mode, &cm);
      if (rc) return rc;
      b[i] += lr * (cp - cm) / 2.0;
    }
  }
  if (gammas_out) std::copy(g.begin(), g.end(), gammas_out);
  if (betas_out) std::copy(b.begin(), b.end(), betas_out);
  res->best_cut_value = best;
  res->best_index = best_idx;
  res->iterations = hist;
  res->converged = conv ? 1 : 0;
  res->evaluations = q->evaluations;
  res->device_ms = q->device_ms;
  res->wall_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  return RQ_OK;
}

int rq_argmax_probability(const rq_state* h, uint64_t* index, double* prob) {
  State* s = (State*)h;
  if (!s) return fail(RQ_ERR_ARGUMENT, "null state");
  return argmax_probability(*s, index, prob);
}

int rq_expectation_zmasks(const rq_state* h, const uint64_t* masks, size_t n, double* out) {
  State* s = (State*)h;
  if (!s || (n && (!masks || !out))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  for (size_t k = 0; k < n; ++k)
    if (s->nq < 64 && (masks[k] >> s->nq)) return fail(RQ_ERR_INVALID_QUBIT, "Z mask names a qubit beyond the state");
  return n ? reduce_zmasks(*s, masks, (int)n, out) : RQ_OK;
}

}  // extern "C"
