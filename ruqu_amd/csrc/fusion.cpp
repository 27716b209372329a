// Host planner of the fused schedule (see fusion.h) and the fused runner.
//
// Planning steps for one segment (a run of unitary gates between
// measurements):
//  1. 1q fusion: consecutive single-qubit gates on a qubit are multiplied
//     into one 2x2 (C = next * fused, the order of optimizer.rs:70), across
//     gates on other qubits; a 2q gate on the qubit flushes it.
//  2. 2q gates become items: CNOT(c,t) = X on t controlled by c; CZ and Rzz =
//     2-bit diagonals; SWAP = dense 4x4.
//  3. Stages: greedy scan; an item joins the current stage when none of its
//     qubits is blocked by an earlier deferred item and its non-diagonal
//     targets fit in the m-qubit tile set; otherwise it is deferred and
//     blocks its qubits.  Items on disjoint qubits commute, so the order on
//     every qubit is preserved.
//  4. Phases inside a stage: the same greedy rule with 4 register qubits.
#include <algorithm>
#include <cstring>
#include <string>
#include <thread>

#include "fusion.h"
#include "gates.h"

namespace rq {

namespace {

struct Item {
  int kind;       // FOP_*
  int q0, q1;     // targets (U1: q0; U2: q0 = row-high, q1; D1: q0; D2: q0 high bit, q1 low bit)
  int ctl;        // control qubit of a U1 (-1: none)
  int mat;        // pool offset
  uint64_t touched() const {
    uint64_t m = 1ull << q0;
    if (kind == FOP_U2 || kind == FOP_D2 || kind == FOP_CZ) m |= 1ull << q1;
    if (ctl >= 0) m |= 1ull << ctl;
    return m;
  }
  uint64_t need() const {  // qubits that must be register qubits
    if (kind == FOP_U1 || kind == FOP_CX) return 1ull << q0;
    if (kind == FOP_U2) return (1ull << q0) | (1ull << q1);
    return 0;
  }
};

inline void mat_mul_2x2(const rq_c64 a[4], const rq_c64 b[4], rq_c64 c[4]) {  // optimizer.rs:12-23
  c[0] = cadd(cmul(a[0], b[0]), cmul(a[1], b[2]));
  c[1] = cadd(cmul(a[0], b[1]), cmul(a[1], b[3]));
  c[2] = cadd(cmul(a[2], b[0]), cmul(a[3], b[2]));
  c[3] = cadd(cmul(a[2], b[1]), cmul(a[3], b[3]));
}

int popc(uint64_t x) { return __builtin_popcountll(x); }

}  // namespace

FusionConfig& fusion_config() {
  static FusionConfig cfg;
  return cfg;
}

int plan_fused(uint32_t nq, const rq_gate* gates, size_t ngates, const FusionConfig& cfg, FusedPlan* plan) {
  plan->stages.clear();
  plan->phases.clear();
  plan->ops.clear();
  plan->pool.clear();
  plan->unitary = 0;
  if (nq < (uint32_t)FUSE_R_BITS) return fail(RQ_ERR_ARGUMENT, "fused schedule needs at least 4 qubits");
  std::vector<rq_c64>& pool = plan->pool;

  // ---- 1. + 2.: items -------------------------------------------------------
  std::vector<Item> items;
  std::vector<int> pending(nq, -1);  // pool offset of the pending 2x2 per qubit
  auto flush = [&](uint32_t q) {
    if (pending[q] >= 0) {
      items.push_back(Item{FOP_U1, (int)q, -1, -1, pending[q]});
      pending[q] = -1;
    }
  };
  for (size_t k = 0; k < ngates; ++k) {
    const rq_gate& g = gates[k];
    if (g.kind == RQ_GATE_BARRIER) continue;
    if (g.kind == RQ_GATE_MEASURE || g.kind == RQ_GATE_RESET)
      return fail(RQ_ERR_ARGUMENT, "plan_fused: segment contains a measurement");
    if (g.q0 >= nq || (is_2q(g.kind) && g.q1 >= nq))
      return fail(RQ_ERR_INVALID_QUBIT, "invalid qubit index " + std::to_string(std::max(g.q0, is_2q(g.kind) ? g.q1 : 0)) +
                                            " for " + std::to_string(nq) + "-qubit system");
    ++plan->unitary;
    if (is_1q(g.kind)) {
      rq_c64 m[4];
      matrix_1q(g, m);
      if (pending[g.q0] >= 0) {
        rq_c64 c[4];
        mat_mul_2x2(m, &pool[pending[g.q0]], c);
        std::memcpy(&pool[pending[g.q0]], c, sizeof(c));
      } else {
        pending[g.q0] = (int)pool.size();
        pool.insert(pool.end(), m, m + 4);
      }
      continue;
    }
    if (g.q0 == g.q1)
      return fail(RQ_ERR_CIRCUIT, "two-qubit gate requires distinct qubits, got " + std::to_string(g.q0) + " and " +
                                      std::to_string(g.q1));
    flush(g.q0);
    flush(g.q1);
    rq_c64 m[16];
    matrix_2q(g, m);
    int off = (int)pool.size();
    switch (g.kind) {
      case RQ_GATE_CNOT:
        items.push_back(Item{FOP_CX, (int)g.q1, -1, (int)g.q0, 0});
        break;
      case RQ_GATE_CZ:
        items.push_back(Item{FOP_CZ, (int)g.q0, (int)g.q1, -1, 0});
        break;
      case RQ_GATE_RZZ: {
        rq_c64 d[4] = {m[0], m[5], m[10], m[15]};
        pool.insert(pool.end(), d, d + 4);
        items.push_back(Item{FOP_D2, (int)g.q0, (int)g.q1, -1, off});
        break;
      }
      default:
        pool.insert(pool.end(), m, m + 16);
        items.push_back(Item{FOP_U2, (int)g.q0, (int)g.q1, -1, off});
    }
  }
  for (uint32_t q = 0; q < nq; ++q) flush(q);
  if (items.empty()) return RQ_OK;

  // ---- 3.: stages -----------------------------------------------------------
  const int m = std::min<int>(std::min(cfg.m, FUSE_MAX_M), (int)nq);
  const int c = std::max(0, std::min(cfg.c, m - 2));  // leaves room for any 2q target pair
  if ((int)nq - m > FUSE_MAX_OUTER) return fail(RQ_ERR_ARGUMENT, "state too large for the fused schedule");
  const uint64_t low = c >= 64 ? ~0ull : ((1ull << c) - 1);
  std::vector<int> remaining(items.size());
  for (size_t k = 0; k < items.size(); ++k) remaining[k] = (int)k;

  // One greedy scan from the tile-qubit seed S0 (items in order; an item
  // joins when its qubits are not blocked and its targets fit the tile).
  auto scan = [&](uint64_t S0, std::vector<int>* ops, std::vector<int>* deferred, size_t window) {
    uint64_t S = S0, blocked = 0;
    int n_ops = 0;
    for (size_t k = 0; k < remaining.size() && k < window; ++k) {
      const int id = remaining[k];
      const Item& it = items[id];
      const uint64_t t = it.touched();
      const uint64_t ns = S | it.need();
      if (!(t & blocked) && popc(ns) <= m) {
        S = ns;
        ++n_ops;
        if (ops) ops->push_back(id);
      } else {
        blocked |= t;
        if (deferred) deferred->push_back(id);
      }
    }
    return std::make_pair(n_ops, S);
  };
  uint64_t rng = 0x9E3779B97F4A7C15ull;  // xorshift64: fixed seed, so plans are deterministic
  auto next_rand = [&]() {
    rng ^= rng << 13;
    rng ^= rng >> 7;
    rng ^= rng << 17;
    return rng;
  };
  while (!remaining.empty()) {
    // Seed search: the stage's tile set is decided by the items that come
    // first, so also try seeds made of the needed qubits of the front items
    // (the first unblocked ones) in random orders, and keep the seed that
    // takes the most items into this stage (ties: the plain scan).  Seeds are
    // compared over the first SEED_WINDOW remaining items, which bounds the
    // planning time.
    uint64_t seed = low;
    if (cfg.seeds > 0 && (int)nq > m) {
      int front[64], nfront = 0;
      uint64_t blocked = 0;
      for (int id : remaining) {
        const uint64_t nd = items[id].need() & ~low;
        for (uint64_t b = nd & ~blocked; b && nfront < 64; b &= b - 1) {
          const int q = __builtin_ctzll(b);
          bool dup = false;
          for (int k = 0; k < nfront; ++k) dup |= front[k] == q;
          if (!dup) front[nfront++] = q;
        }
        blocked |= items[id].touched();
        if (nfront >= 64 || popc(blocked) >= (int)nq) break;
      }
      constexpr size_t SEED_WINDOW = 384;
      int best = scan(low, nullptr, nullptr, SEED_WINDOW).first;
      // The seeds in RNG order, then their scores on host threads (each scan
      // only reads the items), then the first best in seed order: the plan
      // does not depend on the thread count.
      std::vector<uint64_t> cand;
      for (int r = 0; r < cfg.seeds && nfront > m - popc(low); ++r) {
        for (int k = nfront - 1; k > 0; --k) std::swap(front[k], front[next_rand() % (uint64_t)(k + 1)]);
        uint64_t S0 = low;
        for (int k = 0; k < nfront && popc(S0) < m; ++k) S0 |= 1ull << front[k];
        cand.push_back(S0);
      }
      std::vector<int> score(cand.size());
      const unsigned hw = std::thread::hardware_concurrency();
      const size_t nt = std::min<size_t>({hw ? hw : 1u, 16u, (cand.size() + 31) / 32});
      auto work = [&](size_t t) {
        for (size_t i = t; i < cand.size(); i += nt) score[i] = scan(cand[i], nullptr, nullptr, SEED_WINDOW).first;
      };
      if (nt <= 1) {
        work(0);
      } else {
        std::vector<std::thread> pool;
        for (size_t t = 1; t < nt; ++t) pool.emplace_back(work, t);
        work(0);
        for (auto& th : pool) th.join();
      }
      for (size_t i = 0; i < cand.size(); ++i)
        if (score[i] > best) best = score[i], seed = cand[i];
    }
    std::vector<int> ops, deferred;
    uint64_t S = scan(seed, &ops, &deferred, remaining.size()).second;
    // Fill the tile up to m qubits: first the qubits the deferred items need
    // (no cost now, and the next stage may start from a similar set), then the
    // lowest free qubits.
    for (int id : deferred) {
      if (popc(S) >= m) break;
      uint64_t nd = items[id].need() & ~S;
      while (nd && popc(S) < m) {
        uint64_t b = nd & (~nd + 1);
        S |= b;
        nd &= nd - 1;
      }
    }
    for (uint32_t q = 0; q < nq && popc(S) < m; ++q) S |= 1ull << q;

    FStage st{};
    st.m = m;
    int tile_of[64];
    for (int q = 0; q < 64; ++q) tile_of[q] = -1;
    int ti = 0, oi = 0;
    for (uint32_t q = 0; q < nq; ++q) {
      if (S >> q & 1) {
        st.S[ti] = (int)q;
        tile_of[q] = ti++;
      } else {
        st.outer[oi++] = (int)q;
      }
    }
    st.nouter = oi;
    st.phase0 = (int)plan->phases.size();

    // ---- 4.: phases ---------------------------------------------------------
    std::vector<int> prem = ops;
    while (!prem.empty()) {
      uint64_t R = 0, pblocked = 0;
      std::vector<int> pops, pdef;
      for (int id : prem) {
        const Item& it = items[id];
        uint64_t t = it.touched();
        if (t & pblocked) {
          pdef.push_back(id);
          pblocked |= t;
          continue;
        }
        uint64_t nr = R | it.need();
        if (popc(nr) <= FUSE_R_BITS) {
          R = nr;
          pops.push_back(id);
        } else {
          pdef.push_back(id);
          pblocked |= t;
        }
      }
      // Pad the register set to 4 qubits, preferring tile bits >= 6 so that
      // the lanes of a wave stay on tile bits 0..5 (direct HBM access).
      for (int tb = m - 1; tb >= 0 && popc(R) < FUSE_R_BITS; --tb)
        if (tb >= 6 || m - FUSE_R_BITS < 6) R |= 1ull << st.S[tb];
      for (int tb = 0; tb < m && popc(R) < FUSE_R_BITS; ++tb) R |= 1ull << st.S[tb];

      FPhase ph{};
      ph.mf_off = -1;
      int nb = 0;
      int reg_of[64];
      for (int q = 0; q < 64; ++q) reg_of[q] = -1;
      for (int tb = 0; tb < m; ++tb)
        if (R >> st.S[tb] & 1) {
          reg_of[st.S[tb]] = nb;
          ph.regt[nb++] = tb;
        }
      int thr_of[64];
      for (int q = 0; q < 64; ++q) thr_of[q] = -1;
      int nt = 0;
      for (int tb = 0; tb < m; ++tb)
        if (!(R >> st.S[tb] & 1)) {
          thr_of[st.S[tb]] = nt;
          ph.thrt[nt++] = tb;
        }
      // With full waves the first and last phases access HBM in their own
      // mapping.  The lanes hold the lowest non-register tile bits, so the
      // 16 accesses of a thread together cover every line they touch; this
      // measured faster than a pass through the natural mapping and one more
      // LDS transpose, also when a register bit is a low tile bit.
      ph.direct = (m - FUSE_R_BITS) >= 6 ? 1 : 0;
      ph.op_begin = (int)plan->ops.size();
      auto locate = [&](int q, int8_t* loc, int32_t* pos) {
        if (reg_of[q] >= 0) { *loc = LOC_REG; *pos = reg_of[q]; }
        else if (thr_of[q] >= 0) { *loc = LOC_THR; *pos = thr_of[q]; }
        else { *loc = LOC_OUT; *pos = q; }
      };
      for (int id : pops) {
        const Item& it = items[id];
        FOp op{};
        op.kind = it.kind;
        op.mat = it.mat;
        if (it.kind == FOP_U1 || it.kind == FOP_CX) {
          op.t0 = reg_of[it.q0];
          if (it.ctl >= 0) {
            int8_t loc; int32_t pos;
            locate(it.ctl, &loc, &pos);
            if (loc == LOC_REG) op.ctl_reg = 1u << pos;
            else if (loc == LOC_THR) op.ctl_thr = 1u << pos;
            else op.ctl_out = 1ull << pos;
          }
        } else if (it.kind == FOP_U2) {
          op.t0 = reg_of[it.q0];
          op.t1 = reg_of[it.q1];
        } else if (it.kind == FOP_D1) {
          locate(it.q0, &op.dloc0, &op.dpos0);
        } else {  // D2, CZ
          locate(it.q0, &op.dloc0, &op.dpos0);
          locate(it.q1, &op.dloc1, &op.dpos1);
        }
        plan->ops.push_back(op);
      }
      ph.op_end = (int)plan->ops.size();
      plan->phases.push_back(ph);
      prem = pdef;
    }
    st.nphase = (int)plan->phases.size() - st.phase0;
    plan->stages.push_back(st);
    remaining = deferred;
  }
  return RQ_OK;
}

int run_fused(State& s, const rq_gate* gates, size_t ngates, const double* draws, size_t ndraws, size_t* cursor,
              rq_measurement* meas, size_t cap, size_t* produced, FusionStats* st, bool zero_first) {
  const double pass_bytes = 2.0 * (double)s.n * (double)s.amp_bytes();
  // zero_first: the caller wants |0..0> first; the first fused stage makes it
  // in registers, anything else clears the state first
  bool zero = zero_first;
  auto settle = [&]() {
    if (zero) launch_init_zero(s);
    zero = false;
  };
  size_t k = 0;
  while (k < ngates) {
    size_t e = k;
    while (e < ngates && gates[e].kind != RQ_GATE_MEASURE && gates[e].kind != RQ_GATE_RESET) ++e;
    if (e > k) {
      if (s.nq < (uint32_t)FUSE_R_BITS) {
        // Tiny states: the exact per-gate kernels.
        settle();
        for (size_t j = k; j < e; ++j) {
          int has = 0;
          rq_measurement mm;
          int rc = apply_gate_impl(s, gates[j], 0.0, &mm, &has);
          if (rc) return rc;
          if (!is_non_unitary(gates[j].kind)) {
            ++st->unitary;
            ++st->passes;
            st->hbm_bytes += gate_bytes(s, gates[j].kind);
          }
        }
      } else {
        uint64_t unitary = 0, stages = 0;
        int rc = run_fused_segment(s, gates + k, e - k, &unitary, &stages, &zero);
        if (rc) return rc;
        settle();  // a segment without stages leaves the state as it was
        st->unitary += unitary;
        st->passes += stages;
        st->hbm_bytes += pass_bytes * (double)stages;
        s.gate_count += unitary;
      }
    }
    if (e < ngates) {
      settle();
      const rq_gate& g = gates[e];
      if (*cursor >= ndraws) return fail(RQ_ERR_CIRCUIT, "measurement needs a uniform draw but the draw list is exhausted");
      double draw = draws[(*cursor)++];
      if (g.q0 >= s.nq)
        return fail(RQ_ERR_INVALID_QUBIT, "invalid qubit index " + std::to_string(g.q0) + " for " +
                                              std::to_string(s.nq) + "-qubit system");
      if (g.kind == RQ_GATE_MEASURE) {
        rq_measurement mm;
        int rc = measure_state(s, g.q0, draw, &mm);
        if (rc) return rc;
        if (meas && *produced < cap) meas[*produced] = mm;
        ++*produced;
      } else {
        int rc = reset_state(s, g.q0, draw);
        if (rc) return rc;
      }
    }
    k = e + 1;
  }
  settle();
  st->fused = st->unitary > st->passes ? st->unitary - st->passes : 0;
  return RQ_OK;
}

}  // namespace rq

extern "C" {

// Plans gates[0..n) (no measurements) for a state of nq qubits with tile m and
// low-qubit count c and reports the schedule's size.  Host only.
int rq_fusion_plan_info(uint32_t nq, const rq_gate* gates, size_t ngates, int m, int c, uint64_t* stages,
                        uint64_t* phases, uint64_t* ops) {
  rq::FusedPlan plan;
  rq::FusionConfig cfg = rq::fusion_config();  // the current knobs, with this tile shape
  cfg.m = m;
  cfg.c = c;
  int rc = rq::plan_fused(nq, gates, ngates, cfg, &plan);
  if (rc) return rc;
  if (stages) *stages = plan.stages.size();
  if (phases) *phases = plan.phases.size();
  if (ops) *ops = plan.ops.size();
  return RQ_OK;
}

// Runs the fused schedule of gates[0..n) on host amplitudes (rq_c64 or rq_c32
// per precision): the program the kernel would receive, with its data mapping
// and the current tuning knobs.  Host only: used to check the
// planner and the packing without a GPU.
int rq_fusion_emulate(uint32_t nq, int precision, const rq_gate* gates, size_t ngates, int m, int c, void* amps) {
  rq::FusedPlan plan;
  rq::FusionConfig cfg = rq::fusion_config();  // the current knobs, with this tile shape
  cfg.m = m;
  cfg.c = c;
  int rc = rq::plan_fused(nq, gates, ngates, cfg, &plan);
  if (rc) return rc;
  return rq::emulate_packed(plan, precision, rq::fusion_config().mf, amps);
}

// Size of the program the kernel would receive for gates[0..n): op records,
// matrix-core phases and their composed unitaries (current tuning knobs).
// Host only.
int rq_fusion_packed_info(uint32_t nq, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                          uint64_t* records, uint64_t* mf_phases, uint64_t* mf_unitaries) {
  rq::FusedPlan plan;
  rq::FusionConfig cfg = rq::fusion_config();  // the current knobs, with this tile shape
  cfg.m = m;
  cfg.c = c;
  int rc = rq::plan_fused(nq, gates, ngates, cfg, &plan);
  if (rc) return rc;
  return rq::packed_info(plan, precision, rq::fusion_config().mf, records, mf_phases, mf_unitaries);
}

// Record counts of that program by kind (counts has 9 entries).  Host only.
int rq_fusion_packed_kinds(uint32_t nq, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                           uint64_t* counts) {
  if (!counts) return rq::fail(RQ_ERR_ARGUMENT, "null pointer");
  rq::FusedPlan plan;
  rq::FusionConfig cfg = rq::fusion_config();  // the current knobs, with this tile shape
  cfg.m = m;
  cfg.c = c;
  int rc = rq::plan_fused(nq, gates, ngates, cfg, &plan);
  if (rc) return rc;
  return rq::packed_kinds(plan, precision, rq::fusion_config().mf, counts);
}

// The packed records themselves: word x of each and the phase it belongs to
// (-1 for the records of matrix-core phases).  Host only.
int rq_fusion_packed_records(uint32_t nq, int precision, const rq_gate* gates, size_t ngates, int m, int c,
                             uint32_t* words, int32_t* phase, size_t cap, size_t* count) {
  if (!words || !phase || !count) return rq::fail(RQ_ERR_ARGUMENT, "null pointer");
  rq::FusedPlan plan;
  rq::FusionConfig cfg = rq::fusion_config();
  cfg.m = m;
  cfg.c = c;
  int rc = rq::plan_fused(nq, gates, ngates, cfg, &plan);
  if (rc) return rc;
  return rq::packed_records(plan, precision, rq::fusion_config().mf, words, phase, cap, count);
}

// Sets the tile size (m qubits) and the count of low qubits always kept in a
// tile (c) used by the fused schedule of rq_run_circuit.
int rq_fusion_configure(int m, int c) {
  if (m < 4 || m > rq::FUSE_MAX_M || c < 0 || c > m) return rq::fail(RQ_ERR_ARGUMENT, "fusion configuration out of range");
  rq::fusion_config().m = m;
  rq::fusion_config().c = c;
  return RQ_OK;
}

}  // extern "C"
