// Batched QAOA / VQE driver: instance registry, per-round work lists, the
// device-side optimizer updates and the hipGraph of one optimizer iteration.
//
// Optimizer semantics per instance are those of the reference:
//   QAOA plain (qaoa.rs:217-308): evaluate; history; best (> , argmax
//   bitstring); converged when |cost - prev| < 1e-6 (iteration > 0); then
//   gamma_0..gamma_{p-1} and beta_0..beta_{p-1} updated one after another,
//   each from two pi/2-shifted evaluations at the angles updated so far.
//   VQE (vqe.rs:148-216): evaluate; history; best (<, params); converged when
//   |prev - e| < threshold; gradient of every parameter from two shifted
//   evaluations at the same parameters; simultaneous descent update.
// Round 0 of an iteration evaluates every QAOA base circuit and every VQE
// circuit (base and all shifts); round r >= 1 evaluates the two shifted
// circuits of QAOA parameter r-1.  Each round is one launch of the persistent
// kernel; the update kernels run on the device between rounds, so a whole
// iteration is captured once as a hipGraph and replayed.
//
// Shared prefixes (states above 13 qubits): the two circuits shifted at
// gamma_l or beta_l agree with each other, at the angles updated so far, up to
// the input of pass l (pass s applies layer s's cost phase and starts its
// mixer).  Each such instance keeps a running prefix state; in the round of
// gamma_l (beta_l) an advance item runs pass l-1 on it with the current
// angles, and the two shifted circuits start at pass l from it.  The gamma
// rounds and the beta rounds each start the chain from |+> (l = 0).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "batch.h"
#include "internal.h"

namespace rq {

__global__ void k_batch_eval(const BInst* insts, const BItem* items, int nitems, unsigned* counter,
                             const int32_t* done, const BEdge* edges, const BTerm* terms, const double* params,
                             const float* ctab, double* energy, uint32_t* amax);
template <int TQ>
__global__ void k_wide_pass(const BInst* insts, const BItem* witems, const int32_t* wmap, const int32_t* wtile0,
                            const int64_t* wstate, const WidePasses wp, const int32_t* done, const double* params,
                            const float* ctab, float2* states, int pass, BPart* parts);
__global__ void k_wide_energy(const BInst* insts, const BItem* witems, const int32_t* wmap, const int32_t* wtile0,
                              const int64_t* wstate, const int32_t* vmap, const int32_t* done, const BTerm* terms,
                              const float2* states, BPart* parts);
__global__ void k_wide_fold(const BInst* insts, const BItem* witems, const int32_t* wtile0, const int32_t* done,
                            const BPart* parts, double* energy, uint32_t* amax);
__global__ void k_cut_tables(float* ctab, const BInst* insts, const BEdge* edges, const BCutChunk* chunks,
                             uint32_t nchunks);
__global__ void k_maxcut(int n, const BEdge* e, int ne, unsigned long long* out);
__global__ void k_zero_u32(unsigned* p, int n);

namespace {

struct OptState {
  int32_t* iter;
  int32_t* hlen;
  int32_t* conv;
  int32_t* done;
  double* hist;
  double* best;
  uint32_t* best_idx;
  double* best_params;
};

__global__ void k_opt_init(const BInst* insts, int ninst, OptState o) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst) return;
  o.iter[k] = 0;
  o.hlen[k] = 0;
  o.conv[k] = 0;
  o.done[k] = 0;
  o.best[k] = insts[k].kind == BK_QAOA ? -INFINITY : 1.7976931348623157e308;
  o.best_idx[k] = 0;
}

__global__ void k_update0(const BInst* insts, int ninst, const int32_t* out0, const double* energy,
                          const uint32_t* amax, double* params, OptState o, int max_iter, double vqe_lr,
                          double vqe_thr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst || o.done[k]) return;
  const BInst in = insts[k];
  const int it = o.iter[k];
  const double e = energy[out0[k]];
  double* h = o.hist + (size_t)k * max_iter;
  h[it] = e;
  o.hlen[k] = it + 1;
  if (in.kind == BK_QAOA) {
    if (e > o.best[k]) {  // qaoa.rs:240-256; the angles of the best energy are kept as well
      o.best[k] = e;
      o.best_idx[k] = amax[out0[k]];
      for (int j = 0; j < in.nparams; ++j) o.best_params[in.param_off + j] = params[in.param_off + j];
    }
    if (it > 0 && fabs(e - h[it - 1]) < 1e-6) {
      o.conv[k] = 1;
      o.done[k] = 1;
    }
  } else {
    if (e < o.best[k]) {
      o.best[k] = e;
      for (int j = 0; j < in.nparams; ++j) o.best_params[in.param_off + j] = params[in.param_off + j];
    }
    if (it > 0 && fabs(h[it - 1] - e) < vqe_thr) {
      o.conv[k] = 1;
      o.done[k] = 1;
      return;
    }
    for (int j = 0; j < in.nparams; ++j) {
      double g = (energy[out0[k] + 1 + 2 * j] - energy[out0[k] + 2 + 2 * j]) / 2.0;
      params[in.param_off + j] -= vqe_lr * g;
    }
    o.iter[k] = it + 1;
    if (it + 1 >= max_iter) o.done[k] = 1;
  }
}

__global__ void k_update_r(const BInst* insts, int ninst, int r, const int32_t* outr, const double* energy,
                           double* params, OptState o, double qaoa_lr) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst || o.done[k]) return;
  const BInst in = insts[k];
  if (in.kind != BK_QAOA || r > in.nparams) return;
  const int o_ = outr[k] + 2 * (r - 1);
  const double cp = energy[o_], cm = energy[o_ + 1];
  params[in.param_off + r - 1] += qaoa_lr * (cp - cm) / 2.0;  // qaoa.rs:285,296
}

__global__ void k_end_iter(const BInst* insts, int ninst, OptState o, int max_iter) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= ninst || o.done[k] || insts[k].kind != BK_QAOA) return;
  o.iter[k] += 1;
  if (o.iter[k] >= max_iter) o.done[k] = 1;
}

template <class T>
int upload(T** d, const std::vector<T>& h) {
  size_t bytes = std::max<size_t>(h.size(), 1) * sizeof(T);
  RQ_HIP(hipMalloc((void**)d, bytes));
  if (!h.empty()) RQ_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return RQ_OK;
}

}  // namespace

// Launch plan of one round of circuit evaluations.  Small items (n <= 13)
// go to the persistent kernel; wide items are sorted by layer count
// (descending), so the items still active at layer l are a prefix and a
// layer's launch covers exactly their tiles.
struct RoundPlan {
  int small_off = 0, nsmall = 0;  // into all_small
  int wide_off = 0, nwide = 0;    // into all_wide / all_wtile0 / all_wstate / all_wfirst / all_wlast / all_wsrc
  int tile_off = 0, ntiles = 0;   // into all_wmap
  // k_wide_pass launches, pass by pass and, within a pass, one per tile width:
  // the pass, the tile qubits, the number of tiles and the tile list's offset
  // in all_ptile
  std::vector<int> launch_pass, launch_tq, layer_grid, pass_off;
  int vmap_off = 0, nvtiles = 0;  // into all_vmap: the round's VQE tiles
  int64_t state_amps = 0;         // HBM amplitudes needed (prefix states included)
};

// Passes of one wide item: first..last; the first pass (when > 0) reads the
// state at src instead of the item's own slot.  own: the item's state
// (amplitude offset), or -1 for a slot of its own in the round.
struct ItemPass {
  int first = 0, last = -1;
  int64_t src = -1;
  int64_t own = -1;
};

struct Batch {
  std::vector<BInst> insts;
  std::vector<BEdge> edges;
  std::vector<BTerm> terms;
  std::vector<double> params;
  bool ready = false;     // device data uploaded
  bool planned = false;   // host plan built for the current tile width
  bool plan_on_device = false;
  int tq = BATCH_TILE_QUBITS;
  bool tile_exact = false;  // every wide state on tiles of tq qubits (else batch_item_tq)
  int nwg = 0;
  hipStream_t st = nullptr;
  BInst* d_insts = nullptr;
  BEdge* d_edges = nullptr;
  BTerm* d_terms = nullptr;
  double* d_params = nullptr;
  float* d_ctab = nullptr;
  int64_t ctab_total = 0;
  BCutChunk* d_cutchunks = nullptr;
  // plan: rounds 0..nrounds-1 of an optimizer iteration, then the base round
  int nrounds = 0;
  std::vector<RoundPlan> rounds;
  RoundPlan base;
  std::vector<BItem> all_small, all_wide;
  std::vector<int32_t> all_wtile0, all_wmap, all_vmap, all_wfirst, all_wlast, all_ptile;
  std::vector<int64_t> all_wstate, all_wsrc;
  std::vector<int64_t> pfx_off;  // per instance: its prefix state (amplitudes), -1: none
  int64_t prefix_amps = 0;       // prefix states, at the start of the state buffer
  BItem* d_small = nullptr;
  BItem* d_wide = nullptr;
  int32_t* d_wtile0 = nullptr;
  int32_t* d_wmap = nullptr;
  int32_t* d_vmap = nullptr;
  int64_t* d_wstate = nullptr;
  int32_t* d_wfirst = nullptr;
  int32_t* d_wlast = nullptr;
  int64_t* d_wsrc = nullptr;
  int32_t* d_ptile = nullptr;
  float2* d_states = nullptr;
  int64_t states_cap = 0;  // amplitudes
  BPart* d_parts = nullptr;
  int parts_cap = 0;
  int nouts = 0;
  double* d_energy = nullptr;
  uint32_t* d_amax = nullptr;
  unsigned* d_counters = nullptr;
  int32_t* d_out0 = nullptr;
  int32_t* d_outr = nullptr;
  std::vector<int32_t> out0, outr;
  OptState o{};
  int hist_cap = 0;
  double* d_best_params = nullptr;
  int32_t* d_zero_done = nullptr;  // all-zero flags for plain evaluation
  hipGraphExec_t exec = nullptr;
  int graph_iter_cap = -1;
  double last_ms = 0;
};

namespace {

// Dynamic LDS of k_wide_pass: 2^tq amplitudes plus one pad slot per 2^(tq-8).
size_t wide_lds_bytes(int tq) { return ((size_t(1) << tq) + (size_t(1) << 8)) * sizeof(float2); }

double item_cost(const BInst& in) {
  double amps = std::ldexp(1.0, in.n);
  return in.kind == BK_QAOA ? amps * (in.p + 1) : amps * (in.p + in.nterms);
}

// Appends one round's items to the flat plan arrays.  pv[k]: the passes of
// item v[k] when it is wide (last < 0: all of them, 0..p).  The round's own
// states start at amplitude amps0: rounds after round 0 keep them clear of
// the QAOA prefix states (which live from one round to the next), round 0 of
// the VQE chains (which live within round 0); round 0 and plain evaluation
// may overlap the QAOA prefixes (the gamma chain restarts from |+> in round 1).
RoundPlan plan_round(Batch& b, const std::vector<BItem>& v, std::vector<ItemPass> pv = {}, int64_t amps0 = 0) {
  RoundPlan rp;
  pv.resize(v.size());
  std::vector<BItem> small;
  std::vector<std::pair<BItem, ItemPass>> wide;
  for (size_t k = 0; k < v.size(); ++k) {
    if (b.insts[v[k].inst].n <= BATCH_LDS_QUBITS) {
      small.push_back(v[k]);
      continue;
    }
    ItemPass ip = pv[k];
    if (ip.last < 0) ip.last = b.insts[v[k].inst].p;
    wide.emplace_back(v[k], ip);
  }
  std::stable_sort(small.begin(), small.end(), [&](const BItem& x, const BItem& y) {
    return item_cost(b.insts[x.inst]) > item_cost(b.insts[y.inst]);
  });
  std::stable_sort(wide.begin(), wide.end(), [&](const auto& x, const auto& y) {
    const BInst &a = b.insts[x.first.inst], &c = b.insts[y.first.inst];
    return a.p != c.p ? a.p > c.p : a.n > c.n;
  });
  rp.small_off = (int)b.all_small.size();
  rp.nsmall = (int)small.size();
  b.all_small.insert(b.all_small.end(), small.begin(), small.end());
  rp.wide_off = (int)b.all_wide.size();
  rp.nwide = (int)wide.size();
  rp.tile_off = (int)b.all_wmap.size();
  int tiles = 0, maxpass = -1;
  int64_t amps = amps0;
  rp.vmap_off = (int)b.all_vmap.size();
  for (int j = 0; j < rp.nwide; ++j) {
    const BInst& in = b.insts[wide[j].first.inst];
    const ItemPass& ip = wide[j].second;
    const int nt = 1 << (in.n - in.tq);
    b.all_wide.push_back(wide[j].first);
    b.all_wtile0.push_back(tiles);
    b.all_wfirst.push_back(ip.first);
    b.all_wlast.push_back(ip.last);
    if (ip.own >= 0) {
      b.all_wstate.push_back(ip.own);
    } else {
      b.all_wstate.push_back(amps);
      amps += (int64_t)1 << in.n;
    }
    b.all_wsrc.push_back(ip.first > 0 ? ip.src : b.all_wstate.back());
    for (int t = 0; t < nt; ++t) {
      b.all_wmap.push_back(j);
      if (in.kind == BK_VQE && wide[j].first.out >= 0) b.all_vmap.push_back(tiles + t);
    }
    tiles += nt;
    maxpass = std::max(maxpass, ip.last);
  }
  // pass s, tile width w: the tiles of every item on w-qubit tiles whose
  // passes include s, item by item; one launch each (widest first)
  for (int s = 0; s <= maxpass; ++s)
    for (int w = BATCH_MAX_TILE_QUBITS; w >= BATCH_MIN_TILE_QUBITS; --w) {
      const int off = (int)b.all_ptile.size();
      for (int j = 0; j < rp.nwide; ++j) {
        const ItemPass& ip = wide[j].second;
        const BInst& in = b.insts[wide[j].first.inst];
        if (in.tq != w || s < ip.first || s > ip.last) continue;
        const int t0 = b.all_wtile0[rp.wide_off + j], nt = 1 << (in.n - in.tq);
        for (int t = 0; t < nt; ++t) b.all_ptile.push_back(t0 + t);
      }
      if ((int)b.all_ptile.size() == off) continue;
      rp.launch_pass.push_back(s);
      rp.launch_tq.push_back(w);
      rp.layer_grid.push_back((int)b.all_ptile.size() - off);
      rp.pass_off.push_back(off);
    }
  rp.nvtiles = (int)b.all_vmap.size() - rp.vmap_off;
  rp.ntiles = tiles;
  rp.state_amps = amps;
  return rp;
}

// Host part of the setup: cut-table offsets, output slots and the launch
// plan of every round.  No device call.
void build_plan(Batch& b) {
  const int ninst = (int)b.insts.size();
  int nmax = 0;
  for (const BInst& in : b.insts)
    if (in.n > BATCH_LDS_QUBITS) nmax = std::max(nmax, in.n);
  b.tq = std::max(b.tq, (nmax + 1) / 2);  // two tile passes per layer
  for (BInst& in : b.insts)
    in.tq = in.n > BATCH_LDS_QUBITS ? (b.tile_exact ? b.tq : batch_item_tq(in.n, b.tq)) : 0;
  b.ctab_total = 0;
  for (BInst& in : b.insts) {
    if (in.kind == BK_QAOA) {  // C(i) tables for every QAOA state, LDS-resident ones included
      in.ctab_off = b.ctab_total;
      b.ctab_total += (int64_t)1 << in.n;
    } else {
      in.ctab_off = -1;
    }
  }
  int pmax2 = 0;
  for (const BInst& in : b.insts)
    if (in.kind == BK_QAOA) pmax2 = std::max(pmax2, in.nparams);
  b.nrounds = 1 + pmax2;
  // prefix states of the wide QAOA instances with a shift beyond layer 0
  b.pfx_off.assign(ninst, -1);
  b.prefix_amps = 0;
  for (int k = 0; k < ninst; ++k) {
    const BInst& in = b.insts[k];
    if (in.kind == BK_QAOA && in.n > BATCH_LDS_QUBITS && in.p >= 2) {
      b.pfx_off[k] = b.prefix_amps;
      b.prefix_amps += (int64_t)1 << in.n;
    }
  }
  // VQE (simultaneous gradient): the base circuit of a wide instance with
  // L >= 2 layers runs as a chain of single-pass items that keep each pass's
  // output (chain slot s = output of pass s), and the circuits shifted at
  // layer l start at pass l from chain slot l - 1.
  std::vector<int64_t> chain_off(ninst, -1);
  int64_t chain_amps = 0;
  for (int k = 0; k < ninst; ++k) {
    const BInst& in = b.insts[k];
    if (in.kind == BK_VQE && in.n > BATCH_LDS_QUBITS && in.p >= 2) {
      chain_off[k] = chain_amps;
      chain_amps += (int64_t)in.p << in.n;
    }
  }
  std::vector<std::vector<BItem>> rounds(b.nrounds);
  std::vector<std::vector<ItemPass>> rpass(b.nrounds);
  std::vector<BItem> base;
  b.out0.assign(ninst, -1);
  b.outr.assign(ninst, -1);
  int out = 0;
  for (int k = 0; k < ninst; ++k) {
    const BInst& in = b.insts[k];
    b.out0[k] = out;
    std::vector<BItem>& r0 = rounds[0];
    std::vector<ItemPass>& p0 = rpass[0];
    p0.resize(r0.size());
    const int64_t N = (int64_t)1 << in.n;
    const bool chain = chain_off[k] >= 0;
    auto slot = [&](int s) { return chain_off[k] + (int64_t)s * N; };
    if (chain)
      for (int s = 0; s < in.p; ++s) {  // passes 0..L-1 of the base circuit, each kept
        ItemPass c;
        c.first = c.last = s;
        c.src = s > 0 ? slot(s - 1) : -1;
        c.own = slot(s);
        r0.push_back(BItem{k, 0, -1, -1});
        p0.push_back(c);
      }
    ItemPass bp;
    if (chain) {  // the base circuit's last pass, from the chain
      bp.first = in.p;
      bp.src = slot(in.p - 1);
    }
    r0.push_back(BItem{k, 0, -1, out++});
    p0.push_back(bp);
    base.push_back(BItem{k, 0, -1, b.out0[k]});
    if (in.kind == BK_VQE)
      for (int j = 0; j < in.nparams; ++j) {
        const int l = j / in.n;  // parameter (layer l, qubit j % n)
        ItemPass sp;
        if (chain && l >= 1) {
          sp.first = l;
          sp.src = slot(l - 1);
        }
        r0.push_back(BItem{k, 1, j, out++});
        p0.push_back(sp);
        r0.push_back(BItem{k, 2, j, out++});
        p0.push_back(sp);
      }
  }
  for (int k = 0; k < ninst; ++k) {
    const BInst& in = b.insts[k];
    if (in.kind != BK_QAOA) continue;
    b.outr[k] = out;
    for (int j = 0; j < in.nparams; ++j) {
      const int l = j < in.p ? j : j - in.p;  // gamma_l, then beta_l
      const bool shared = b.pfx_off[k] >= 0 && l >= 1;
      std::vector<BItem>& rv = rounds[1 + j];
      std::vector<ItemPass>& pv = rpass[1 + j];
      pv.resize(rv.size());
      if (shared) {  // advance the prefix: pass l-1 at the current angles
        ItemPass adv;
        adv.first = adv.last = l - 1;
        adv.src = adv.own = b.pfx_off[k];
        rv.push_back(BItem{k, 0, -1, -1});
        pv.push_back(adv);
      }
      ItemPass sh;
      if (shared) {
        sh.first = l;
        sh.src = b.pfx_off[k];
      }
      rv.push_back(BItem{k, 1, j, out++});
      pv.push_back(sh);
      rv.push_back(BItem{k, 2, j, out++});
      pv.push_back(sh);
    }
  }
  b.nouts = out;
  b.all_small.clear();
  b.all_wide.clear();
  b.all_wtile0.clear();
  b.all_wmap.clear();
  b.all_vmap.clear();
  b.all_wstate.clear();
  b.all_wfirst.clear();
  b.all_wlast.clear();
  b.all_wsrc.clear();
  b.all_ptile.clear();
  b.rounds.clear();
  for (int r = 0; r < b.nrounds; ++r) b.rounds.push_back(plan_round(b, rounds[r], rpass[r], r > 0 ? b.prefix_amps : chain_amps));
  b.base = plan_round(b, base);
  b.planned = true;
  b.plan_on_device = false;
}

template <class T>
void free_dev(T*& p) {
  if (p) (void)hipFree(p);
  p = nullptr;
}

// Uploads the launch plan and sizes the wide-state and partial buffers.
int upload_plan(Batch& b) {
  if (b.d_insts)  // the instances' tile widths change with the plan
    RQ_HIP(hipMemcpy(b.d_insts, b.insts.data(), b.insts.size() * sizeof(BInst), hipMemcpyHostToDevice));
  free_dev(b.d_small);
  free_dev(b.d_wide);
  free_dev(b.d_wtile0);
  free_dev(b.d_wmap);
  free_dev(b.d_vmap);
  free_dev(b.d_wstate);
  free_dev(b.d_wfirst);
  free_dev(b.d_wlast);
  free_dev(b.d_wsrc);
  free_dev(b.d_ptile);
  int rc = upload(&b.d_small, b.all_small);
  if (!rc) rc = upload(&b.d_wide, b.all_wide);
  if (!rc) rc = upload(&b.d_wtile0, b.all_wtile0);
  if (!rc) rc = upload(&b.d_wmap, b.all_wmap);
  if (!rc) rc = upload(&b.d_vmap, b.all_vmap);
  if (!rc) rc = upload(&b.d_wstate, b.all_wstate);
  if (!rc) rc = upload(&b.d_wfirst, b.all_wfirst);
  if (!rc) rc = upload(&b.d_wlast, b.all_wlast);
  if (!rc) rc = upload(&b.d_wsrc, b.all_wsrc);
  if (!rc) rc = upload(&b.d_ptile, b.all_ptile);
  if (rc) return rc;
  int64_t amps = b.base.state_amps;
  int parts = b.base.ntiles;
  for (const RoundPlan& rp : b.rounds) {
    amps = std::max(amps, rp.state_amps);
    parts = std::max(parts, rp.ntiles);
  }
  if (amps > b.states_cap) {
    free_dev(b.d_states);
    hipError_t e = hipMalloc((void**)&b.d_states, (size_t)amps * sizeof(float2));
    if (e != hipSuccess) {
      (void)hipGetLastError();
      b.states_cap = 0;
      return fail(RQ_ERR_ALLOC, "batch states: need " + std::to_string((size_t)amps * 8) + " bytes");
    }
    b.states_cap = amps;
  }
  if (parts > b.parts_cap) {
    free_dev(b.d_parts);
    RQ_HIP(hipMalloc((void**)&b.d_parts, (size_t)std::max(parts, 1) * sizeof(BPart)));
    b.parts_cap = parts;
  }
  b.plan_on_device = true;
  return RQ_OK;
}

int finalize(Batch& b) {
  if (!b.planned) build_plan(b);
  if (b.ready) return b.plan_on_device ? RQ_OK : upload_plan(b);
  if (b.insts.empty()) return fail(RQ_ERR_ARGUMENT, "empty batch");
  int dev = 0, cus = 0;
  RQ_HIP(hipGetDevice(&dev));
  RQ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  b.nwg = 2 * std::max(cus, 1);
  RQ_HIP(hipStreamCreateWithFlags(&b.st, hipStreamNonBlocking));
  RQ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_batch_eval), hipFuncAttributeMaxDynamicSharedMemorySize,
                             (1 << BATCH_LDS_QUBITS) * 8));
  RQ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_wide_pass<13>),
                             hipFuncAttributeMaxDynamicSharedMemorySize, wide_lds_bytes(13)));
  const int ninst = (int)b.insts.size();
  int rc = upload(&b.d_insts, b.insts);
  if (!rc) rc = upload(&b.d_edges, b.edges);
  if (!rc) rc = upload(&b.d_terms, b.terms);
  if (!rc) rc = upload(&b.d_params, b.params);
  if (rc) return rc;
  RQ_HIP(hipMalloc((void**)&b.d_best_params, std::max<size_t>(b.params.size(), 1) * sizeof(double)));
  RQ_HIP(hipMalloc((void**)&b.d_ctab, std::max<int64_t>(b.ctab_total, 1) * sizeof(float)));
  {  // every cut table in one launch
    std::vector<BCutChunk> chunks;
    for (size_t k = 0; k < b.insts.size(); ++k)
      if (b.insts[k].ctab_off >= 0) {
        const int64_t N = (int64_t)1 << b.insts[k].n;
        for (int64_t c = 0; c * BATCH_CUT_CHUNK < N; ++c) chunks.push_back(BCutChunk{(int32_t)k, (uint32_t)c});
      }
    if (chunks.size() >= ((size_t)1 << 32)) return fail(RQ_ERR_ARGUMENT, "batch cut tables too large");
    if (!chunks.empty()) {
      rc = upload(&b.d_cutchunks, chunks);
      if (rc) return rc;
      hipLaunchKernelGGL(k_cut_tables, grid_blocks(chunks.size()), dim3(BATCH_THREADS), 0, b.st, b.d_ctab, b.d_insts,
                         b.d_edges, b.d_cutchunks, (uint32_t)chunks.size());
    }
  }
  rc = upload(&b.d_out0, b.out0);
  if (!rc) rc = upload(&b.d_outr, b.outr);
  if (!rc) rc = upload_plan(b);
  if (rc) return rc;
  RQ_HIP(hipMalloc((void**)&b.d_energy, std::max(b.nouts, 1) * sizeof(double)));
  RQ_HIP(hipMalloc((void**)&b.d_amax, std::max(b.nouts, 1) * sizeof(uint32_t)));
  RQ_HIP(hipMalloc((void**)&b.d_counters, (b.nrounds + 1) * sizeof(unsigned)));
  std::vector<int32_t> zeros(ninst, 0);
  rc = upload(&b.d_zero_done, zeros);
  if (rc) return rc;
  RQ_HIP(hipMalloc((void**)&b.o.iter, ninst * sizeof(int32_t)));
  RQ_HIP(hipMalloc((void**)&b.o.hlen, ninst * sizeof(int32_t)));
  RQ_HIP(hipMalloc((void**)&b.o.conv, ninst * sizeof(int32_t)));
  RQ_HIP(hipMalloc((void**)&b.o.done, ninst * sizeof(int32_t)));
  RQ_HIP(hipMalloc((void**)&b.o.best, ninst * sizeof(double)));
  RQ_HIP(hipMalloc((void**)&b.o.best_idx, ninst * sizeof(uint32_t)));
  b.o.best_params = b.d_best_params;
  RQ_HIP(hipStreamSynchronize(b.st));
  b.ready = true;
  return RQ_OK;
}

// One round: the persistent kernel over the small items, then the layer
// passes, tile partials and folds of the wide items.
void enqueue_round(Batch& b, const RoundPlan& rp, unsigned* counter, const int32_t* done) {
  if (rp.nsmall) {
    const unsigned grid = (unsigned)std::min(b.nwg, rp.nsmall);
    hipLaunchKernelGGL(k_batch_eval, dim3(grid), dim3(BATCH_THREADS), (1 << BATCH_LDS_QUBITS) * 8, b.st, b.d_insts,
                       b.d_small + rp.small_off, rp.nsmall, counter, done, b.d_edges, b.d_terms, b.d_params, b.d_ctab,
                       b.d_energy, b.d_amax);
  }
  if (!rp.nwide) return;
  const BItem* wi = b.d_wide + rp.wide_off;
  const int32_t* wm = b.d_wmap + rp.tile_off;
  const int32_t* w0 = b.d_wtile0 + rp.wide_off;
  const int64_t* ws = b.d_wstate + rp.wide_off;
  for (size_t li = 0; li < rp.layer_grid.size(); ++li) {
    const int ps = rp.launch_pass[li], tq = rp.launch_tq[li];
    const size_t lds = wide_lds_bytes(tq);
    const dim3 g(rp.layer_grid[li]), t(BATCH_THREADS);
    const WidePasses wp{b.d_ptile + rp.pass_off[li], b.d_wfirst + rp.wide_off, b.d_wlast + rp.wide_off,
                        b.d_wsrc + rp.wide_off};
    switch (tq) {
      case 9: hipLaunchKernelGGL(k_wide_pass<9>, g, t, lds, b.st, b.d_insts, wi, wm, w0, ws, wp, done, b.d_params,
                                 b.d_ctab, b.d_states, ps, b.d_parts); break;
      case 10: hipLaunchKernelGGL(k_wide_pass<10>, g, t, lds, b.st, b.d_insts, wi, wm, w0, ws, wp, done, b.d_params,
                                  b.d_ctab, b.d_states, ps, b.d_parts); break;
      case 11: hipLaunchKernelGGL(k_wide_pass<11>, g, t, lds, b.st, b.d_insts, wi, wm, w0, ws, wp, done, b.d_params,
                                  b.d_ctab, b.d_states, ps, b.d_parts); break;
      case 12: hipLaunchKernelGGL(k_wide_pass<12>, g, t, lds, b.st, b.d_insts, wi, wm, w0, ws, wp, done, b.d_params,
                                  b.d_ctab, b.d_states, ps, b.d_parts); break;
      default: hipLaunchKernelGGL(k_wide_pass<13>, g, t, lds, b.st, b.d_insts, wi, wm, w0, ws, wp, done, b.d_params,
                                  b.d_ctab, b.d_states, ps, b.d_parts); break;
    }
  }
  if (rp.nvtiles)
    hipLaunchKernelGGL(k_wide_energy, dim3(rp.nvtiles), dim3(BATCH_THREADS), 0, b.st, b.d_insts, wi, wm, w0, ws,
                       b.d_vmap + rp.vmap_off, done, b.d_terms, b.d_states, b.d_parts);
  hipLaunchKernelGGL(k_wide_fold, dim3(rp.nwide), dim3(BATCH_THREADS), 0, b.st, b.d_insts, wi, w0, done, b.d_parts,
                     b.d_energy, b.d_amax);
}

void enqueue_iteration(Batch& b, int max_iter, double qaoa_lr, double vqe_lr, double vqe_thr) {
  const int ninst = (int)b.insts.size();
  const dim3 g1((ninst + 255) / 256), t1(256);
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, b.st, b.d_counters, b.nrounds + 1);
  for (int r = 0; r < b.nrounds; ++r) {
    enqueue_round(b, b.rounds[r], b.d_counters + r, b.o.done);
    if (r == 0)
      hipLaunchKernelGGL(k_update0, g1, t1, 0, b.st, b.d_insts, ninst, b.d_out0, b.d_energy, b.d_amax, b.d_params,
                         b.o, max_iter, vqe_lr, vqe_thr);
    else
      hipLaunchKernelGGL(k_update_r, g1, t1, 0, b.st, b.d_insts, ninst, r, b.d_outr, b.d_energy, b.d_params, b.o,
                         qaoa_lr);
  }
  hipLaunchKernelGGL(k_end_iter, g1, t1, 0, b.st, b.d_insts, ninst, b.o, max_iter);
}

void free_batch(Batch& b) {
  if (b.st) (void)hipStreamSynchronize(b.st);
  if (b.exec) (void)hipGraphExecDestroy(b.exec);
  void* ptrs[] = {b.d_insts,  b.d_edges,  b.d_terms,  b.d_params, b.d_ctab,     b.d_small,   b.d_wide,
                  b.d_wtile0, b.d_wmap,   b.d_vmap,   b.d_wstate, b.d_states, b.d_parts,   b.d_energy,  b.d_amax,
                  b.d_counters, b.d_out0, b.d_outr,   b.o.iter,   b.o.hlen,     b.o.conv,    b.o.done,
                  b.o.hist,   b.o.best,   b.o.best_idx, b.d_best_params, b.d_zero_done, b.d_cutchunks,
                  b.d_wfirst, b.d_wlast,  b.d_wsrc,   b.d_ptile};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (b.st) (void)hipStreamDestroy(b.st);
}

}  // namespace
}  // namespace rq

using namespace rq;

extern "C" {

int rq_batch_create(rq_batch** out) {
  if (!out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  *out = (rq_batch*)new Batch();
  return RQ_OK;
}

void rq_batch_destroy(rq_batch* h) {
  Batch* b = (Batch*)h;
  if (!b) return;
  free_batch(*b);
  delete b;
}

int rq_batch_add_qaoa(rq_batch* h, const rq_graph* g, uint32_t p, const double* init_gammas, const double* init_betas,
                      uint32_t* id) {
  Batch* b = (Batch*)h;
  if (!b || !g || (g->num_edges && (!g->ei || !g->ej || !g->w))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (b->ready) return fail(RQ_ERR_ARGUMENT, "batch already finalized");
  if (g->num_nodes < 2 || g->num_nodes > (uint32_t)BATCH_MAX_QUBITS)
    return fail(RQ_ERR_QUBIT_LIMIT, "batched instances take 2.." + std::to_string(BATCH_MAX_QUBITS) + " qubits");
  for (uint32_t e = 0; e < g->num_edges; ++e)
    if (g->ei[e] >= g->num_nodes || g->ej[e] >= g->num_nodes)
      return fail(RQ_ERR_INVALID_QUBIT, "node index out of range");
  BInst in{};
  in.kind = BK_QAOA;
  in.n = (int)g->num_nodes;
  in.p = (int)p;
  in.nparams = 2 * (int)p;
  in.nterms = (int)g->num_edges;
  in.term_off = (int)b->edges.size();
  in.param_off = (int)b->params.size();
  in.wsum = 0.0;
  for (uint32_t e = 0; e < g->num_edges; ++e) {
    b->edges.push_back(BEdge{g->ei[e], g->ej[e], g->w[e]});
    in.wsum += g->w[e];
  }
  for (uint32_t l = 0; l < p; ++l) b->params.push_back(init_gammas ? init_gammas[l] : 0.5);  // qaoa.rs:221
  for (uint32_t l = 0; l < p; ++l) b->params.push_back(init_betas ? init_betas[l] : 0.5);    // qaoa.rs:222
  if (id) *id = (uint32_t)b->insts.size();
  b->insts.push_back(in);
  b->planned = false;
  return RQ_OK;
}

int rq_batch_add_vqe(rq_batch* h, uint32_t n, uint32_t layers, const double* coeffs, const uint64_t* offsets,
                     const rq_pauli_factor* ops, size_t nterms, const double* init_params, uint32_t* id) {
  Batch* b = (Batch*)h;
  if (!b || (nterms && (!coeffs || !offsets))) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (b->ready) return fail(RQ_ERR_ARGUMENT, "batch already finalized");
  if (n < 2 || n > (uint32_t)BATCH_MAX_QUBITS)
    return fail(RQ_ERR_QUBIT_LIMIT, "batched VQE instances take 2.." + std::to_string(BATCH_MAX_QUBITS) + " qubits");
  BInst in{};
  in.kind = BK_VQE;
  in.n = (int)n;
  in.p = (int)layers;
  in.nparams = (int)(n * layers);
  in.nterms = (int)nterms;
  in.term_off = (int)b->terms.size();
  in.param_off = (int)b->params.size();
  for (size_t k = 0; k < nterms; ++k) {
    BTerm t{0, 0, 0, 0, coeffs[k]};
    for (uint64_t f = offsets[k]; f < offsets[k + 1]; ++f) {  // canonical form as state.rs:370-392
      uint32_t q = ops[f].qubit;
      int op = ops[f].op;
      if (op < 0 || op > 3) return fail(RQ_ERR_ARGUMENT, "unknown Pauli operator");
      if (op == RQ_PAULI_I) continue;
      if (q >= n) return fail(RQ_ERR_INVALID_QUBIT, "Pauli factor on a qubit beyond the instance");
      if (op == RQ_PAULI_X || op == RQ_PAULI_Y) t.x ^= 1ull << q;
      if (op == RQ_PAULI_Y || op == RQ_PAULI_Z) t.zy ^= 1ull << q;
      if (op == RQ_PAULI_Y) t.ny += 1;
    }
    t.ny &= 3;
    b->terms.push_back(t);
  }
  for (int j = 0; j < in.nparams; ++j) b->params.push_back(init_params ? init_params[j] : 0.1);  // vqe.rs:152
  if (id) *id = (uint32_t)b->insts.size();
  b->insts.push_back(in);
  b->planned = false;
  return RQ_OK;
}

// Energy of every instance at its current parameters (one persistent launch).
int rq_batch_evaluate(rq_batch* h, double* energies, uint64_t* argmax) {
  Batch* b = (Batch*)h;
  if (!b || !energies) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = finalize(*b);
  if (rc) return rc;
  hipEvent_t e0, e1;
  RQ_HIP(hipEventCreate(&e0));
  RQ_HIP(hipEventCreate(&e1));
  hipLaunchKernelGGL(k_zero_u32, dim3(1), dim3(256), 0, b->st, b->d_counters, b->nrounds + 1);
  RQ_HIP(hipEventRecord(e0, b->st));
  enqueue_round(*b, b->base, b->d_counters + b->nrounds, b->d_zero_done);
  RQ_HIP(hipEventRecord(e1, b->st));
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipStreamSynchronize(b->st));
  float ms = 0;
  RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  b->last_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  std::vector<double> en(b->nouts);
  std::vector<uint32_t> am(b->nouts);
  RQ_HIP(hipMemcpy(en.data(), b->d_energy, b->nouts * sizeof(double), hipMemcpyDeviceToHost));
  RQ_HIP(hipMemcpy(am.data(), b->d_amax, b->nouts * sizeof(uint32_t), hipMemcpyDeviceToHost));
  for (size_t k = 0; k < b->insts.size(); ++k) {
    energies[k] = en[b->out0[k]];
    if (argmax) argmax[k] = am[b->out0[k]];
  }
  return RQ_OK;
}

// Runs max_iter optimizer iterations for every instance (use_graph: capture
// one iteration as a hipGraph and replay it).
int rq_batch_optimize(rq_batch* h, uint32_t max_iter, double qaoa_lr, double vqe_lr, double vqe_threshold,
                      int use_graph, double* device_ms) {
  Batch* b = (Batch*)h;
  if (!b) return fail(RQ_ERR_ARGUMENT, "null batch");
  int rc = finalize(*b);
  if (rc) return rc;
  const int ninst = (int)b->insts.size();
  if ((int)max_iter > b->hist_cap) {
    if (b->o.hist) RQ_HIP(hipFree(b->o.hist));
    RQ_HIP(hipMalloc((void**)&b->o.hist, (size_t)ninst * std::max<uint32_t>(max_iter, 1) * sizeof(double)));
    b->hist_cap = (int)max_iter;
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
  }
  hipLaunchKernelGGL(k_opt_init, dim3((ninst + 255) / 256), dim3(256), 0, b->st, b->d_insts, ninst, b->o);
  RQ_HIP(hipGetLastError());
  hipEvent_t e0, e1;
  RQ_HIP(hipEventCreate(&e0));
  RQ_HIP(hipEventCreate(&e1));
  if (use_graph && (!b->exec || b->graph_iter_cap != (int)max_iter)) {
    if (b->exec) {
      (void)hipGraphExecDestroy(b->exec);
      b->exec = nullptr;
    }
    hipGraph_t graph;
    RQ_HIP(hipStreamSynchronize(b->st));
    RQ_HIP(hipStreamBeginCapture(b->st, hipStreamCaptureModeThreadLocal));
    enqueue_iteration(*b, (int)max_iter, qaoa_lr, vqe_lr, vqe_threshold);
    RQ_HIP(hipStreamEndCapture(b->st, &graph));
    RQ_HIP(hipGraphInstantiate(&b->exec, graph, nullptr, nullptr, 0));
    RQ_HIP(hipGraphDestroy(graph));
    b->graph_iter_cap = (int)max_iter;
  }
  RQ_HIP(hipEventRecord(e0, b->st));
  for (uint32_t it = 0; it < max_iter; ++it) {
    if (use_graph) RQ_HIP(hipGraphLaunch(b->exec, b->st));
    else enqueue_iteration(*b, (int)max_iter, qaoa_lr, vqe_lr, vqe_threshold);
  }
  RQ_HIP(hipEventRecord(e1, b->st));
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipStreamSynchronize(b->st));
  float ms = 0;
  RQ_HIP(hipEventElapsedTime(&ms, e0, e1));
  b->last_ms = ms;
  if (device_ms) *device_ms = ms;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return RQ_OK;
}

// Per-instance results after rq_batch_optimize.  params receives the final
// parameters (QAOA: gammas then betas), best_params the best VQE parameters.
int rq_batch_result(rq_batch* h, uint32_t id, double* best, uint64_t* best_index, uint32_t* iterations,
                    int* converged, double* params, double* best_params, double* history) {
  Batch* b = (Batch*)h;
  if (!b || !b->ready || id >= b->insts.size()) return fail(RQ_ERR_ARGUMENT, "bad batch or instance id");
  const BInst& in = b->insts[id];
  int32_t hl = 0, cv = 0;
  uint32_t bi = 0;
  double bv = 0;
  if (b->o.hist) {
    RQ_HIP(hipMemcpy(&hl, b->o.hlen + id, sizeof(hl), hipMemcpyDeviceToHost));
    RQ_HIP(hipMemcpy(&cv, b->o.conv + id, sizeof(cv), hipMemcpyDeviceToHost));
    RQ_HIP(hipMemcpy(&bi, b->o.best_idx + id, sizeof(bi), hipMemcpyDeviceToHost));
    RQ_HIP(hipMemcpy(&bv, b->o.best + id, sizeof(bv), hipMemcpyDeviceToHost));
    if (history && hl > 0)
      RQ_HIP(hipMemcpy(history, b->o.hist + (size_t)id * b->hist_cap, hl * sizeof(double), hipMemcpyDeviceToHost));
  }
  if (best) *best = bv;
  if (best_index) *best_index = bi;
  if (iterations) *iterations = (uint32_t)hl;
  if (converged) *converged = cv;
  if (params)
    RQ_HIP(hipMemcpy(params, b->d_params + in.param_off, in.nparams * sizeof(double), hipMemcpyDeviceToHost));
  if (best_params && b->o.hist)
    RQ_HIP(hipMemcpy(best_params, b->d_best_params + in.param_off, in.nparams * sizeof(double),
                     hipMemcpyDeviceToHost));
  return RQ_OK;
}

// Overwrites the parameters of one instance (e.g. a warm start) before the
// next evaluate/optimize.
int rq_batch_set_params(rq_batch* h, uint32_t id, const double* params) {
  Batch* b = (Batch*)h;
  if (!b || !params || id >= b->insts.size()) return fail(RQ_ERR_ARGUMENT, "bad batch or instance id");
  const BInst& in = b->insts[id];
  std::copy(params, params + in.nparams, b->params.begin() + in.param_off);
  if (b->ready)
    RQ_HIP(hipMemcpy(b->d_params + in.param_off, params, in.nparams * sizeof(double), hipMemcpyHostToDevice));
  return RQ_OK;
}

int rq_batch_set_all_params(rq_batch* h, const double* params, size_t count) {
  Batch* b = (Batch*)h;
  if (!b || (!params && count)) return fail(RQ_ERR_ARGUMENT, "null pointer");
  if (count != b->params.size())
    return fail(RQ_ERR_ARGUMENT, "expected " + std::to_string(b->params.size()) + " parameters, got " +
                                     std::to_string(count));
  std::copy(params, params + count, b->params.begin());
  if (b->ready && count)
    RQ_HIP(hipMemcpy(b->d_params, params, count * sizeof(double), hipMemcpyHostToDevice));
  return RQ_OK;
}

namespace {

int set_tile(Batch* b, int tq, bool exact) {
  if (!b) return fail(RQ_ERR_ARGUMENT, "null batch");
  int nmax = 0;
  for (const BInst& in : b->insts)
    if (in.n > BATCH_LDS_QUBITS) nmax = std::max(nmax, in.n);
  if (tq < BATCH_MIN_TILE_QUBITS || tq > BATCH_MAX_TILE_QUBITS || 2 * tq < nmax)
    return fail(RQ_ERR_ARGUMENT, "tile qubits must be in [max(" + std::to_string(BATCH_MIN_TILE_QUBITS) +
                                     ", ceil(n_max/2)), " + std::to_string(BATCH_MAX_TILE_QUBITS) + "]");
  if (tq == b->tq && exact == b->tile_exact) return RQ_OK;
  if (b->st) RQ_HIP(hipStreamSynchronize(b->st));
  b->tq = tq;
  b->tile_exact = exact;
  b->planned = false;
  if (b->exec) {
    (void)hipGraphExecDestroy(b->exec);
    b->exec = nullptr;
  }
  return RQ_OK;
}

}  // namespace

// LDS tile width of the wide (n > 13) states: 2^tq amplitudes per
// workgroup, widened per state (batch_item_tq) up to 13 qubits where the
// second tiling would otherwise read runs shorter than a 128-byte line.
// Needs ceil(n_max / 2) <= tq <= 13 (two passes per layer).
int rq_batch_set_tile(rq_batch* h, int tq) { return set_tile((Batch*)h, tq, false); }

// The same with every wide state on tiles of exactly tq qubits.
int rq_batch_set_tile_exact(rq_batch* h, int tq) { return set_tile((Batch*)h, tq, true); }

int rq_batch_get_tile(const rq_batch* h) {
  const Batch* b = (const Batch*)h;
  return b ? b->tq : -1;
}

// Host-only (no device call): the launch plan of one round (round < 0: the
// plain evaluation, else round r of an optimizer iteration), as the device
// receives it.  info[0..7] = nsmall, nwide, ntiles, nlaunches (k_wide_pass
// launches: one per pass and tile width), nrounds, nouts, tq (the batch's
// base tile), nvtiles; info64[0..1] = HBM amplitudes of the round's wide
// states, cut-table floats; pass_grid = tiles per launch.  Null arrays are
// skipped (call once for the counts).
int rq_batch_plan_export(rq_batch* h, int round, int32_t* info, int64_t* info64, void* small_items,
                         void* wide_items, int32_t* wtile0, int64_t* wstate, int32_t* wmap, int32_t* pass_grid,
                         int32_t* vmap) {
  Batch* b = (Batch*)h;
  if (!b || b->insts.empty()) return fail(RQ_ERR_ARGUMENT, "empty batch");
  if (!b->planned) build_plan(*b);
  if (round >= b->nrounds) return fail(RQ_ERR_ARGUMENT, "round out of range");
  const RoundPlan& rp = round < 0 ? b->base : b->rounds[round];
  if (info) {
    const int32_t v[8] = {rp.nsmall, rp.nwide, rp.ntiles, (int32_t)rp.layer_grid.size(), b->nrounds, b->nouts,
                          b->tq, rp.nvtiles};
    std::memcpy(info, v, sizeof(v));
  }
  if (info64) {
    info64[0] = rp.state_amps;
    info64[1] = b->ctab_total;
  }
  if (small_items) std::memcpy(small_items, b->all_small.data() + rp.small_off, rp.nsmall * sizeof(BItem));
  if (wide_items) std::memcpy(wide_items, b->all_wide.data() + rp.wide_off, rp.nwide * sizeof(BItem));
  if (wtile0) std::memcpy(wtile0, b->all_wtile0.data() + rp.wide_off, rp.nwide * sizeof(int32_t));
  if (wstate) std::memcpy(wstate, b->all_wstate.data() + rp.wide_off, rp.nwide * sizeof(int64_t));
  if (wmap) std::memcpy(wmap, b->all_wmap.data() + rp.tile_off, rp.ntiles * sizeof(int32_t));
  if (pass_grid) std::copy(rp.layer_grid.begin(), rp.layer_grid.end(), pass_grid);
  if (vmap) std::memcpy(vmap, b->all_vmap.data() + rp.vmap_off, rp.nvtiles * sizeof(int32_t));
  return RQ_OK;
}

// Host-only: the k_wide_pass launches of one round: the pass and the tile
// qubits of each.
int rq_batch_plan_launches(rq_batch* h, int round, int32_t* pass, int32_t* tile_qubits) {
  Batch* b = (Batch*)h;
  if (!b || b->insts.empty()) return fail(RQ_ERR_ARGUMENT, "empty batch");
  if (!b->planned) build_plan(*b);
  if (round >= b->nrounds) return fail(RQ_ERR_ARGUMENT, "round out of range");
  const RoundPlan& rp = round < 0 ? b->base : b->rounds[round];
  if (pass) std::copy(rp.launch_pass.begin(), rp.launch_pass.end(), pass);
  if (tile_qubits) std::copy(rp.launch_tq.begin(), rp.launch_tq.end(), tile_qubits);
  return RQ_OK;
}

// Host-only: the pass structure of one round's wide items: first / last pass
// and first-pass source state per item, and the tile list of every launch
// (concatenated; launch i starts at pass_off[i]).
int rq_batch_plan_passes(rq_batch* h, int round, int32_t* wfirst, int32_t* wlast, int64_t* wsrc, int32_t* ptile,
                         int32_t* pass_off) {
  Batch* b = (Batch*)h;
  if (!b || b->insts.empty()) return fail(RQ_ERR_ARGUMENT, "empty batch");
  if (!b->planned) build_plan(*b);
  if (round >= b->nrounds) return fail(RQ_ERR_ARGUMENT, "round out of range");
  const RoundPlan& rp = round < 0 ? b->base : b->rounds[round];
  if (wfirst) std::memcpy(wfirst, b->all_wfirst.data() + rp.wide_off, rp.nwide * sizeof(int32_t));
  if (wlast) std::memcpy(wlast, b->all_wlast.data() + rp.wide_off, rp.nwide * sizeof(int32_t));
  if (wsrc) std::memcpy(wsrc, b->all_wsrc.data() + rp.wide_off, rp.nwide * sizeof(int64_t));
  int64_t total = 0;
  for (int g : rp.layer_grid) total += g;
  if (ptile && total)
    std::memcpy(ptile, b->all_ptile.data() + (rp.pass_off.empty() ? 0 : rp.pass_off[0]), total * sizeof(int32_t));
  if (pass_off)
    for (size_t s = 0; s < rp.pass_off.size(); ++s) pass_off[s] = rp.pass_off[s] - rp.pass_off[0];
  return RQ_OK;
}

// Host-only: instance (48 B, with cut-table offsets), edge (16 B),
// Pauli-term (32 B) and parameter records in device layout.
int rq_batch_data_export(rq_batch* h, void* insts, void* edges, void* terms, double* params, size_t* counts) {
  Batch* b = (Batch*)h;
  if (!b) return fail(RQ_ERR_ARGUMENT, "null batch");
  if (!b->insts.empty() && !b->planned) build_plan(*b);
  if (counts) {
    counts[0] = b->insts.size();
    counts[1] = b->edges.size();
    counts[2] = b->terms.size();
    counts[3] = b->params.size();
  }
  if (insts) std::memcpy(insts, b->insts.data(), b->insts.size() * sizeof(BInst));
  if (edges) std::memcpy(edges, b->edges.data(), b->edges.size() * sizeof(BEdge));
  if (terms) std::memcpy(terms, b->terms.data(), b->terms.size() * sizeof(BTerm));
  if (params) std::memcpy(params, b->params.data(), b->params.size() * sizeof(double));
  return RQ_OK;
}

// All instances at once after rq_batch_optimize: best energy, best bitstring
// index, history length, converged flag per instance; params / best_params
// hold every instance's parameters back to back in instance order.
int rq_batch_results(rq_batch* h, double* best, uint64_t* best_index, uint32_t* iterations, int* converged,
                     double* params, double* best_params) {
  Batch* b = (Batch*)h;
  if (!b || !b->ready || !b->o.hist) return fail(RQ_ERR_ARGUMENT, "batch has not been optimized");
  const int ninst = (int)b->insts.size();
  std::vector<int32_t> hl(ninst), cv(ninst);
  std::vector<uint32_t> bi(ninst);
  if (best) RQ_HIP(hipMemcpy(best, b->o.best, ninst * sizeof(double), hipMemcpyDeviceToHost));
  if (best_index) {
    RQ_HIP(hipMemcpy(bi.data(), b->o.best_idx, ninst * sizeof(uint32_t), hipMemcpyDeviceToHost));
    for (int k = 0; k < ninst; ++k) best_index[k] = bi[k];
  }
  if (iterations) {
    RQ_HIP(hipMemcpy(hl.data(), b->o.hlen, ninst * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int k = 0; k < ninst; ++k) iterations[k] = (uint32_t)hl[k];
  }
  if (converged) {
    RQ_HIP(hipMemcpy(cv.data(), b->o.conv, ninst * sizeof(int32_t), hipMemcpyDeviceToHost));
    for (int k = 0; k < ninst; ++k) converged[k] = cv[k];
  }
  if (params) RQ_HIP(hipMemcpy(params, b->d_params, b->params.size() * sizeof(double), hipMemcpyDeviceToHost));
  if (best_params)
    RQ_HIP(hipMemcpy(best_params, b->d_best_params, b->params.size() * sizeof(double), hipMemcpyDeviceToHost));
  return RQ_OK;
}

// Exact maximum cut of every QAOA instance (brute force over all 2^n
// bitstrings on the device, FP64 cut weights); 0 for VQE instances.
int rq_batch_maxcut(rq_batch* h, double* out) {
  Batch* b = (Batch*)h;
  if (!b || !out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = finalize(*b);
  if (rc) return rc;
  const int ninst = (int)b->insts.size();
  unsigned long long* d = nullptr;
  RQ_HIP(hipMalloc((void**)&d, std::max(ninst, 1) * sizeof(unsigned long long)));
  hipLaunchKernelGGL(k_zero_u32, dim3((2 * ninst + 255) / 256), dim3(256), 0, b->st, (unsigned*)d, 2 * ninst);
  for (int k = 0; k < ninst; ++k) {
    const BInst& in = b->insts[k];
    if (in.kind != BK_QAOA) continue;
    const unsigned grid = (unsigned)std::min<uint64_t>(1024, ((1ull << in.n) + 255) / 256);
    hipLaunchKernelGGL(k_maxcut, dim3(grid), dim3(BATCH_THREADS), 0, b->st, in.n, b->d_edges + in.term_off,
                       in.nterms, d + k);
  }
  RQ_HIP(hipGetLastError());
  std::vector<unsigned long long> hv(ninst);
  RQ_HIP(hipMemcpyAsync(hv.data(), d, ninst * sizeof(unsigned long long), hipMemcpyDeviceToHost, b->st));
  RQ_HIP(hipStreamSynchronize(b->st));
  (void)hipFree(d);
  for (int k = 0; k < ninst; ++k) std::memcpy(&out[k], &hv[k], sizeof(double));  // non-negative: bits order = value order
  return RQ_OK;
}

uint32_t rq_batch_size(const rq_batch* h) { return h ? (uint32_t)((const Batch*)h)->insts.size() : 0; }

double rq_batch_last_ms(const rq_batch* h) { return h ? ((const Batch*)h)->last_ms : 0.0; }

}  // extern "C"
