// Fused stage kernel: one workgroup per tile of 2^m amplitudes, every gate of
// the stage applied in registers, one HBM read and one HBM write per
// amplitude per stage (layout and planning in fusion.h / fusion.cpp).
//
// Thread t of a phase holds the 16 amplitudes whose tile index is
//   tile(k, t) = sum_b bit_b(k) << regt[b]  |  sum_j bit_j(t) << thrt[j],
// k = 0..15 over the 4 register qubits.  Phases change the register qubits
// by writing the tile to LDS and reading it back with the next mapping.
// When a phase keeps its register qubits off tile bits 0..5, the 64 lanes of
// a wave cover tile bits 0..5, i.e. runs of contiguous HBM, and the phase
// reads (first phase) or writes (last phase) HBM directly.
//
// Arithmetic: FP contraction is allowed here (fused mode), so results differ
// from the per-gate kernels at the rounding level.
#include <cstring>
#include <utility>

#include "fusion.h"

#pragma clang fp contract(fast)

namespace rq {
namespace {

template <class T>
struct alignas(2 * sizeof(T)) Cx {  // one 8- / 16-byte LDS access per amplitude
  T re, im;
};

template <class T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> m, Cx<T> a) {
  return Cx<T>{m.re * a.re - m.im * a.im, m.re * a.im + m.im * a.re};
}
template <class T>
__device__ __forceinline__ Cx<T> cmadd(Cx<T> m0, Cx<T> a0, Cx<T> m1, Cx<T> a1) {
  return Cx<T>{m0.re * a0.re - m0.im * a0.im + m1.re * a1.re - m1.im * a1.im,
               m0.re * a0.im + m0.im * a0.re + m1.re * a1.im + m1.im * a1.re};
}

template <class T>
struct Vec2;
template <>
struct Vec2<float> {
  using type = float2;
};
template <>
struct Vec2<double> {
  using type = double2;
};

// LDS byte address of tile index t.  The XOR swizzle folds index bits 5..9
// (f32; 4..7 for f64) into the bank bits, so the lanes of a wave hit distinct
// banks whichever tile bits they hold.  It is linear over bits: the address of
// (thread part | register part) is the XOR of the two parts' addresses, so a
// phase computes its per-thread part once and each register adds one XOR.
template <class T>
__device__ __forceinline__ uint32_t lds_addr(uint32_t t) {
  if constexpr (sizeof(T) == 4) return (t ^ ((t >> 5) & 31u)) * 8u;
  else return (t ^ ((t >> 4) & 15u)) * 16u;
}

template <class T>
__device__ __forceinline__ void load_amp(const T* __restrict__ a, uint64_t g, Cx<T>& v) {
  using V = typename Vec2<T>::type;
  V x = reinterpret_cast<const V*>(a)[g];
  v.re = x.x;
  v.im = x.y;
}
template <class T>
__device__ __forceinline__ void store_amp(T* __restrict__ a, uint64_t g, const Cx<T>& v) {
  using V = typename Vec2<T>::type;
  V x;
  x.x = v.re;
  x.y = v.im;
  reinterpret_cast<V*>(a)[g] = x;
}

// 2x2 on register bit B (CTL: only pairs whose k satisfies ctl_reg).
template <int B, bool CTL, class T>
__device__ __forceinline__ void u1(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp, uint32_t ctl_reg) {
  const Cx<T> m00 = mp[0], m01 = mp[1], m10 = mp[2], m11 = mp[3];
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    if (CTL && (k & ctl_reg) != ctl_reg) continue;
    Cx<T> x = v[k], y = v[k1];
    v[k] = cmadd(m00, x, m01, y);
    v[k1] = cmadd(m10, x, m11, y);
  }
}

// 4x4 on register bits (B0 = q1, row-high bit; B1 = q2, row-low bit).
template <class T>
__device__ __forceinline__ Cx<T> row4(const Cx<T>* __restrict__ mp, int r, Cx<T> x0, Cx<T> x1, Cx<T> x2, Cx<T> x3) {
  Cx<T> s = cmadd(mp[r * 4 + 0], x0, mp[r * 4 + 1], x1);
  Cx<T> s2 = cmadd(mp[r * 4 + 2], x2, mp[r * 4 + 3], x3);
  return Cx<T>{s.re + s2.re, s.im + s2.im};
}

template <int B0, int B1, int K, class T>
__device__ __forceinline__ void u2_one(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp) {
  if constexpr ((K & ((1 << B0) | (1 << B1))) == 0) {
    constexpr int i0 = K, i1 = K | (1 << B1), i2 = K | (1 << B0), i3 = K | (1 << B0) | (1 << B1);
    const Cx<T> x0 = v[i0], x1 = v[i1], x2 = v[i2], x3 = v[i3];
    v[i0] = row4(mp, 0, x0, x1, x2, x3);
    v[i1] = row4(mp, 1, x0, x1, x2, x3);
    v[i2] = row4(mp, 2, x0, x1, x2, x3);
    v[i3] = row4(mp, 3, x0, x1, x2, x3);
  }
}

template <int B0, int B1, class T, int... K>
__device__ __forceinline__ void u2_all(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp,
                                       std::integer_sequence<int, K...>) {
  (u2_one<B0, B1, K>(v, mp), ...);
}

template <int B0, int B1, class T>
__device__ __forceinline__ void u2(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp) {
  u2_all<B0, B1>(v, mp, std::make_integer_sequence<int, FUSE_R>{});
}

// Controlled swap on register bit B (CNOT target).
template <int B, class T>
__device__ __forceinline__ void cx(Cx<T> (&v)[FUSE_R], uint32_t ctl_reg) {
#pragma unroll
  for (int k = 0; k < FUSE_R; ++k) {
    if (k & (1 << B)) continue;
    const int k1 = k | (1 << B);
    if ((k & ctl_reg) != ctl_reg) continue;
    const Cx<T> x = v[k];
    v[k] = v[k1];
    v[k1] = x;
  }
}

template <class T>
__device__ __forceinline__ void dispatch_cx(Cx<T> (&v)[FUSE_R], int b, uint32_t ctl_reg) {
  switch (b) {
    case 0: cx<0>(v, ctl_reg); break;
    case 1: cx<1>(v, ctl_reg); break;
    case 2: cx<2>(v, ctl_reg); break;
    default: cx<3>(v, ctl_reg); break;
  }
}

template <bool CTL, class T>
__device__ __forceinline__ void dispatch_u1(Cx<T> (&v)[FUSE_R], int b, const Cx<T>* mp, uint32_t ctl_reg) {
  switch (b) {
    case 0: u1<0, CTL>(v, mp, ctl_reg); break;
    case 1: u1<1, CTL>(v, mp, ctl_reg); break;
    case 2: u1<2, CTL>(v, mp, ctl_reg); break;
    default: u1<3, CTL>(v, mp, ctl_reg); break;
  }
}

template <class T>
__device__ __forceinline__ void dispatch_u2(Cx<T> (&v)[FUSE_R], int b0, int b1, const Cx<T>* mp) {
  switch (b0 * 4 + b1) {
    case 1: u2<0, 1>(v, mp); break;
    case 2: u2<0, 2>(v, mp); break;
    case 3: u2<0, 3>(v, mp); break;
    case 4: u2<1, 0>(v, mp); break;
    case 6: u2<1, 2>(v, mp); break;
    case 7: u2<1, 3>(v, mp); break;
    case 8: u2<2, 0>(v, mp); break;
    case 9: u2<2, 1>(v, mp); break;
    case 11: u2<2, 3>(v, mp); break;
    case 12: u2<3, 0>(v, mp); break;
    case 13: u2<3, 1>(v, mp); break;
    default: u2<3, 2>(v, mp); break;
  }
}

// Register-bit pattern: bit k of the result is bit `pos` of k.
__device__ __forceinline__ uint32_t reg_pattern(int pos) {
  return (uint32_t)((0xFF00F0F0CCCCAAAAull >> (16 * pos)) & 0xFFFFull);
}

// A diagonal index bit of register k: ((kmask >> k) & 1) when the bit is a
// register bit, else the per-thread / per-tile constant cbit.
struct DBit {
  uint32_t kmask;  // register pattern (0 unless the bit is a register bit)
  uint32_t cbit;   // thread or outer bit value (0 for register bits)
  bool reg;
};
__device__ __forceinline__ DBit dbit(int loc, int pos, uint32_t tid, uint64_t outer) {
  DBit d;
  d.reg = loc == LOC_REG;
  d.kmask = d.reg ? reg_pattern(pos) : 0u;
  d.cbit = loc == LOC_THR ? ((tid >> pos) & 1u) : (loc == LOC_OUT ? (uint32_t)((outer >> pos) & 1ull) : 0u);
  return d;
}

template <class T>
__device__ __forceinline__ Cx<T> neg_if(Cx<T> a, bool c) {
  return c ? Cx<T>{-a.re, -a.im} : a;
}

template <class T>
__device__ __forceinline__ void apply_cz(Cx<T> (&v)[FUSE_R], const DBit& b0, const DBit& b1) {
  if (b0.reg && b1.reg) {
    const uint32_t m = b0.kmask & b1.kmask;  // uniform
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k)
      if ((m >> k) & 1u) v[k] = Cx<T>{-v[k].re, -v[k].im};
  } else if (b0.reg || b1.reg) {
    const uint32_t m = b0.kmask | b1.kmask;
    const bool c = (b0.cbit | b1.cbit) != 0u;  // the non-register bit (per thread / tile)
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k)
      if ((m >> k) & 1u) v[k] = neg_if(v[k], c);
  } else {
    const bool c = (b0.cbit & b1.cbit) != 0u;
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = neg_if(v[k], c);
  }
}

template <class T>
__device__ __forceinline__ void apply_d1(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp, const DBit& b0) {
  if (b0.reg) {
    const Cx<T> f0 = mp[0], f1 = mp[1];
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(((b0.kmask >> k) & 1u) ? f1 : f0, v[k]);
  } else {
    const Cx<T> f = b0.cbit ? mp[1] : mp[0];
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(f, v[k]);
  }
}

template <class T>
__device__ __forceinline__ void apply_d2(Cx<T> (&v)[FUSE_R], const Cx<T>* __restrict__ mp, const DBit& b0,
                                         const DBit& b1) {
  if (b0.reg && b1.reg) {
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(mp[((b0.kmask >> k) & 1u) * 2 + ((b1.kmask >> k) & 1u)], v[k]);
  } else if (b0.reg) {
    const Cx<T> f0 = mp[b1.cbit], f1 = mp[2 + b1.cbit];
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(((b0.kmask >> k) & 1u) ? f1 : f0, v[k]);
  } else if (b1.reg) {
    const Cx<T> f0 = mp[b0.cbit * 2], f1 = mp[b0.cbit * 2 + 1];
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(((b1.kmask >> k) & 1u) ? f1 : f0, v[k]);
  } else {
    const Cx<T> f = mp[b0.cbit * 2 + b1.cbit];
#pragma unroll
    for (int k = 0; k < FUSE_R; ++k) v[k] = cmul(f, v[k]);
  }
}

template <class T>
__device__ __forceinline__ void apply_ops(Cx<T> (&v)[FUSE_R], const FOp* __restrict__ ops, int b, int e,
                                          const Cx<T>* __restrict__ pool, uint32_t tid, uint64_t outer) {
  for (int o = b; o < e; ++o) {
    const FOp& op = ops[o];
    const Cx<T>* mp = pool + op.mat;
    switch (op.kind) {
      case FOP_U1:
        if (op.ctl_reg == 0u && op.ctl_thr == 0u && op.ctl_out == 0ull) {
          dispatch_u1<false>(v, op.t0, mp, 0u);
        } else if ((tid & op.ctl_thr) == op.ctl_thr && (outer & op.ctl_out) == op.ctl_out) {
          dispatch_u1<true>(v, op.t0, mp, op.ctl_reg);
        }
        break;
      case FOP_U2:
        dispatch_u2(v, op.t0, op.t1, mp);
        break;
      case FOP_CX:
        if ((tid & op.ctl_thr) == op.ctl_thr && (outer & op.ctl_out) == op.ctl_out) dispatch_cx(v, op.t0, op.ctl_reg);
        break;
      case FOP_CZ:
        apply_cz(v, dbit(op.dloc0, op.dpos0, tid, outer), dbit(op.dloc1, op.dpos1, tid, outer));
        break;
      case FOP_D1:
        apply_d1(v, mp, dbit(op.dloc0, op.dpos0, tid, outer));
        break;
      default:  // FOP_D2
        apply_d2(v, mp, dbit(op.dloc0, op.dpos0, tid, outer), dbit(op.dloc1, op.dpos1, tid, outer));
        break;
    }
  }
}

// Offsets of one phase mapping: thread part and per-register-bit parts of the
// tile index (treg) and of the global index (greg).
struct Map {
  uint32_t tthr;
  uint64_t gthr;
  uint32_t treg[FUSE_R_BITS];
  uint64_t greg[FUSE_R_BITS];
};

template <class T>
__device__ __forceinline__ Cx<T> lds_ld(const Cx<T>* lds, uint32_t byte) {
  return *reinterpret_cast<const Cx<T>*>(reinterpret_cast<const unsigned char*>(lds) + byte);
}
template <class T>
__device__ __forceinline__ void lds_st(Cx<T>* lds, uint32_t byte, const Cx<T>& x) {
  *reinterpret_cast<Cx<T>*>(reinterpret_cast<unsigned char*>(lds) + byte) = x;
}

__device__ __forceinline__ Map make_map(const FStage& st, const int* regt, const int* thrt, uint32_t tid,
                                        uint64_t outer) {
  Map mp;
  mp.tthr = 0;
  mp.gthr = outer;
  const int nt = st.m - FUSE_R_BITS;
  for (int j = 0; j < nt; ++j) {
    uint32_t b = (tid >> j) & 1u;
    mp.tthr |= b << thrt[j];
    mp.gthr |= (uint64_t)b << st.S[thrt[j]];
  }
#pragma unroll
  for (int b = 0; b < FUSE_R_BITS; ++b) {
    mp.treg[b] = 1u << regt[b];
    mp.greg[b] = 1ull << st.S[regt[b]];
  }
  return mp;
}

// The natural mapping: thread bits = tile bits 0..m-5, register bits = the
// top 4 tile bits.
__device__ __forceinline__ Map make_natural_map(const FStage& st, uint32_t tid, uint64_t outer) {
  Map mp;
  const int nt = st.m - FUSE_R_BITS;
  mp.tthr = tid;
  mp.gthr = outer;
  for (int j = 0; j < nt; ++j) mp.gthr |= (uint64_t)((tid >> j) & 1u) << st.S[j];
#pragma unroll
  for (int b = 0; b < FUSE_R_BITS; ++b) {
    mp.treg[b] = 1u << (nt + b);
    mp.greg[b] = 1ull << st.S[nt + b];
  }
  return mp;
}

template <int K>
__device__ __forceinline__ uint32_t treg_of(const Map& mp) {
  return ((K & 1) ? mp.treg[0] : 0u) | ((K & 2) ? mp.treg[1] : 0u) | ((K & 4) ? mp.treg[2] : 0u) |
         ((K & 8) ? mp.treg[3] : 0u);
}
template <int K>
__device__ __forceinline__ uint64_t greg_of(const Map& mp) {
  return ((K & 1) ? mp.greg[0] : 0ull) | ((K & 2) ? mp.greg[1] : 0ull) | ((K & 4) ? mp.greg[2] : 0ull) |
         ((K & 8) ? mp.greg[3] : 0ull);
}

template <int K, class T>
struct Unroll {
  static __device__ __forceinline__ void gather_g(Cx<T> (&v)[FUSE_R], const T* a, const Map& mp) {
    Unroll<K - 1, T>::gather_g(v, a, mp);
    load_amp(a, mp.gthr | greg_of<K - 1>(mp), v[K - 1]);
  }
  static __device__ __forceinline__ void scatter_g(const Cx<T> (&v)[FUSE_R], T* a, const Map& mp) {
    Unroll<K - 1, T>::scatter_g(v, a, mp);
    store_amp(a, mp.gthr | greg_of<K - 1>(mp), v[K - 1]);
  }
  static __device__ __forceinline__ void gather_l(Cx<T> (&v)[FUSE_R], const Cx<T>* lds, const Map& mp,
                                                  uint32_t lt) {
    Unroll<K - 1, T>::gather_l(v, lds, mp, lt);
    v[K - 1] = lds_ld(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)));
  }
  static __device__ __forceinline__ void scatter_l(const Cx<T> (&v)[FUSE_R], Cx<T>* lds, const Map& mp,
                                                   uint32_t lt) {
    Unroll<K - 1, T>::scatter_l(v, lds, mp, lt);
    lds_st(lds, lt ^ lds_addr<T>(treg_of<K - 1>(mp)), v[K - 1]);
  }
};
template <class T>
struct Unroll<0, T> {
  static __device__ __forceinline__ void gather_g(Cx<T> (&)[FUSE_R], const T*, const Map&) {}
  static __device__ __forceinline__ void scatter_g(const Cx<T> (&)[FUSE_R], T*, const Map&) {}
  static __device__ __forceinline__ void gather_l(Cx<T> (&)[FUSE_R], const Cx<T>*, const Map&, uint32_t) {}
  static __device__ __forceinline__ void scatter_l(const Cx<T> (&)[FUSE_R], Cx<T>*, const Map&, uint32_t) {}
};

template <class T>
__global__ __launch_bounds__(256) void k_fused_stage(T* __restrict__ a, const FStage* __restrict__ stp,
                                                      const FPhase* __restrict__ phases, const FOp* __restrict__ ops,
                                                      const Cx<T>* __restrict__ pool) {
  extern __shared__ unsigned char smem_raw[];
  Cx<T>* lds = reinterpret_cast<Cx<T>*>(smem_raw);
  const FStage& st = *stp;
  const uint32_t tid = threadIdx.x;
  uint64_t outer = 0;
  {
    const uint64_t tile = blockIdx.x;
    for (int j = 0; j < st.nouter; ++j) outer |= ((tile >> j) & 1ull) << st.outer[j];
  }
  Cx<T> v[FUSE_R];
  const int np = st.nphase;
  for (int p = 0; p < np; ++p) {
    const FPhase& ph = phases[st.phase0 + p];
    const Map mp = make_map(st, ph.regt, ph.thrt, tid, outer);
    if (p == 0) {
      if (ph.direct) {
        Unroll<FUSE_R, T>::gather_g(v, a, mp);
      } else {
        const Map nat = make_natural_map(st, tid, outer);
        Unroll<FUSE_R, T>::gather_g(v, a, nat);
        Unroll<FUSE_R, T>::scatter_l(v, lds, nat, lds_addr<T>(nat.tthr));
        __syncthreads();
        Unroll<FUSE_R, T>::gather_l(v, lds, mp, lds_addr<T>(mp.tthr));
        __syncthreads();
      }
    } else {
      Unroll<FUSE_R, T>::gather_l(v, lds, mp, lds_addr<T>(mp.tthr));
      __syncthreads();
    }
    apply_ops(v, ops, ph.op_begin, ph.op_end, pool, tid, outer);
    if (p == np - 1) {
      if (ph.direct) {
        Unroll<FUSE_R, T>::scatter_g(v, a, mp);
      } else {
        const Map nat = make_natural_map(st, tid, outer);
        Unroll<FUSE_R, T>::scatter_l(v, lds, mp, lds_addr<T>(mp.tthr));
        __syncthreads();
        Unroll<FUSE_R, T>::gather_l(v, lds, nat, lds_addr<T>(nat.tthr));
        Unroll<FUSE_R, T>::scatter_g(v, a, nat);
      }
    } else {
      Unroll<FUSE_R, T>::scatter_l(v, lds, mp, lds_addr<T>(mp.tthr));
      __syncthreads();
    }
  }
}

template <class T>
int launch_plan_t(State& s, const FusedPlan& plan) {
  const size_t nst = plan.stages.size(), nph = plan.phases.size(), nop = plan.ops.size(), npool = plan.pool.size();
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_st = 0, o_ph = al(nst * sizeof(FStage)), o_op = o_ph + al(nph * sizeof(FPhase)),
               o_pool = o_op + al(nop * sizeof(FOp)), total = o_pool + al(npool * sizeof(Cx<T>));
  if (total > s.prog_cap) {
    if (s.prog) {
      RQ_HIP(hipStreamSynchronize(s.stream));
      RQ_HIP(hipFree(s.prog));
      s.prog = nullptr;
      s.prog_cap = 0;
    }
    size_t cap = total < (1u << 20) ? (1u << 20) : total * 2;
    RQ_HIP(hipMalloc(&s.prog, cap));
    s.prog_cap = cap;
  }
  std::vector<unsigned char> host(total, 0);
  std::memcpy(host.data() + o_st, plan.stages.data(), nst * sizeof(FStage));
  std::memcpy(host.data() + o_ph, plan.phases.data(), nph * sizeof(FPhase));
  std::memcpy(host.data() + o_op, plan.ops.data(), nop * sizeof(FOp));
  Cx<T>* pp = reinterpret_cast<Cx<T>*>(host.data() + o_pool);
  for (size_t k = 0; k < npool; ++k) pp[k] = Cx<T>{(T)plan.pool[k].re, (T)plan.pool[k].im};
  // The host buffer is pageable: wait for the copy before it goes out of scope.
  RQ_HIP(hipMemcpyAsync(s.prog, host.data(), total, hipMemcpyHostToDevice, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  static bool attr_set = false;
  if (!attr_set) {
    RQ_HIP(hipFuncSetAttribute(reinterpret_cast<const void*>(&k_fused_stage<T>),
                               hipFuncAttributeMaxDynamicSharedMemorySize, 160 * 1024));
    attr_set = true;
  }
  unsigned char* base = reinterpret_cast<unsigned char*>(s.prog);
  for (size_t i = 0; i < nst; ++i) {
    const FStage& st = plan.stages[i];
    const unsigned threads = 1u << (st.m - FUSE_R_BITS);
    const uint64_t tiles = 1ull << st.nouter;
    const size_t lds = (size_t(1) << st.m) * sizeof(Cx<T>);
    hipLaunchKernelGGL(k_fused_stage<T>, dim3((unsigned)tiles), dim3(threads), lds, s.stream, (T*)s.amps,
                       reinterpret_cast<const FStage*>(base + o_st + i * sizeof(FStage)),
                       reinterpret_cast<const FPhase*>(base + o_ph), reinterpret_cast<const FOp*>(base + o_op),
                       reinterpret_cast<const Cx<T>*>(base + o_pool));
  }
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

}  // namespace

int launch_fused_plan(State& s, const FusedPlan& plan) {
  if (plan.stages.empty()) return RQ_OK;
  for (const FStage& st : plan.stages)
    if (st.m < FUSE_R_BITS || st.m > FUSE_MAX_M || st.nouter > 31 ||
        (size_t(1) << st.m) * s.amp_bytes() > 160 * 1024)
      return fail(RQ_ERR_ARGUMENT, "fused stage shape out of range");
  if (s.prec == RQ_C64) return launch_plan_t<double>(s, plan);
  return launch_plan_t<float>(s, plan);
}

}  // namespace rq
