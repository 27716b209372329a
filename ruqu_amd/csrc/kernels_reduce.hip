// Reductions over the state: masked norms (measurement, P(q=1)), Pauli-string
// expectations, all <Z_q> at once, inner products, probabilities and the
// sum-tree shot sampler.  Every accumulation is FP64 (also for f32 states):
// per-thread partials, a wave-shuffle + LDS block reduction into one partial
// per block, then one block folds the partials in a fixed order, so results
// are deterministic run to run.
#include <cstring>
#include <mutex>

#include "internal.h"

#pragma clang fp contract(off)

namespace rq {
namespace {

constexpr int BLOCK = 256;
constexpr int WAVE = 64;
constexpr int MAX_RED_BLOCKS = 2048;

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_down(v, off, WAVE);
  return v;
}

// Block-wide sum; result valid in thread 0.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  v = wave_sum(v);
  int lane = threadIdx.x & (WAVE - 1), w = threadIdx.x / WAVE;
  __syncthreads();
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < BLOCK / WAVE; ++k) r += sh[k];
  return r;
}

template <class V>
__device__ __forceinline__ double nsq(V x) {
  double re = (double)x.x, im = (double)x.y;
  return re * re + im * im;
}

// sum_{i : (i & mask) == want} |a_i|^2
template <class V>
__global__ __launch_bounds__(BLOCK) void k_masked_normsq(const V* __restrict__ a, uint64_t n, uint64_t mask,
                                                         uint64_t want, double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  double acc = 0.0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK)
    if ((i & mask) == want) acc += nsq(a[i]);
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// Fold `nparts` partials per term (term-major) into out[term].
__global__ __launch_bounds__(BLOCK) void k_fold(const double* __restrict__ part, int nparts, double* __restrict__ out) {
  __shared__ double sh[BLOCK / WAVE];
  const double* p = part + (size_t)blockIdx.x * nparts;
  double acc = 0.0;
  for (int k = threadIdx.x; k < nparts; k += BLOCK) acc += p[k];
  double r = block_sum(acc, sh);
  if (threadIdx.x == 0) out[blockIdx.x] = r;
}

struct PauliDev {
  uint64_t x, zy;
  uint32_t ny, pad;
};

constexpr int MAX_TERMS_PER_LAUNCH = 32;
struct PauliBatch {
  PauliDev t[MAX_TERMS_PER_LAUNCH];
};

// Re( conj(a_{i^x}) * phase(i) * a_i ), phase(i) = i^ny * (-1)^popcount(i & zy)
// (state.rs:366-396), one FP64 partial per (term, block).
template <class V>
__global__ __launch_bounds__(BLOCK) void k_pauli(const V* __restrict__ a, uint64_t n, PauliBatch b, int nterms,
                                                 double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  for (int t = 0; t < nterms; ++t) {
    const PauliDev pd = b.t[t];
    double acc = 0.0;
    for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
      V ai = a[i];
      V aj = a[i ^ pd.x];
      // phase = i^ny * s
      double s = (__popcll(i & pd.zy) & 1) ? -1.0 : 1.0;
      double pr, pim;
      switch (pd.ny & 3) {
        case 0: pr = s; pim = 0.0; break;
        case 1: pr = 0.0; pim = s; break;
        case 2: pr = -s; pim = 0.0; break;
        default: pr = 0.0; pim = -s; break;
      }
      double cr = (double)aj.x, ci = -(double)aj.y;  // conj(a_j)
      double tr = cr * pr - ci * pim, ti = cr * pim + ci * pr;
      acc += tr * (double)ai.x - ti * (double)ai.y;
    }
    double r = block_sum(acc, sh);
    if (threadIdx.x == 0) part[(size_t)t * gridDim.x + blockIdx.x] = r;
  }
}

// Per block: total norm (slot 0) and P(q=1) for q < nq (slots 1..nq).
template <class V>
__global__ __launch_bounds__(BLOCK) void k_z_all(const V* __restrict__ a, uint64_t n, uint32_t nq,
                                                 double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  double tot = 0.0, p1[40];
#pragma unroll
  for (int q = 0; q < 40; ++q) p1[q] = 0.0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    double v = nsq(a[i]);
    tot += v;
#pragma unroll
    for (int q = 0; q < 40; ++q)
      if ((i >> q) & 1) p1[q] += v;
  }
  double r = block_sum(tot, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
#pragma unroll
  for (uint32_t q = 0; q < 40; ++q) {
    if (q < nq) {  // nq is uniform: every thread takes the same branch
      double rq_ = block_sum(p1[q], sh);
      if (threadIdx.x == 0) part[(size_t)(q + 1) * gridDim.x + blockIdx.x] = rq_;
    }
  }
}

// sum conj(a_i) * b_i (state.rs:430-433): real part then imaginary part.
template <class V>
__global__ __launch_bounds__(BLOCK) void k_inner(const V* __restrict__ a, const V* __restrict__ b, uint64_t n,
                                                 double* __restrict__ part) {
  __shared__ double sh[BLOCK / WAVE];
  double re = 0.0, im = 0.0;
  for (uint64_t i = flat_block() * BLOCK + threadIdx.x; i < n; i += (uint64_t)gridDim.x * BLOCK) {
    V x = a[i], y = b[i];
    double xr = x.x, xi = -(double)x.y, yr = y.x, yi = y.y;
    re += xr * yr - xi * yi;
    im += xr * yi + xi * yr;
  }
  double r = block_sum(re, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
  double r2 = block_sum(im, sh);
  if (threadIdx.x == 0) part[gridDim.x + blockIdx.x] = r2;
}

// P(qubit q = 0) of a complex-f32 state as QuantumStateF32::measure computes
// it (mixed_precision.rs:480-486): p0 = 0f32, then p0 += norm_sq(a_i) in f32,
// in index order over the indices with bit q clear.  That chain of 2^(n-1)
// rounded additions is reproduced bit for bit by one workgroup walking the
// terms in chunks of P0_CHUNK:
//  * while the running sum s is a normal float in the binade [2^E, 2^(E+1)),
//    every float there is a multiple of u = 2^(E-23), so rounding s + x to
//    nearest gives s + r(x) u with r(x) = x/u rounded to the nearest integer,
//    whatever s is, unless x/u lies exactly halfway (then round-to-even looks
//    at s).  If no term of the chunk is such a tie and s/u + sum r < 2^24
//    (every partial sum stays inside the binade; all terms are >= 0), the
//    chunk's chain equals s + (sum r) u: an integer sum done in parallel;
//  * otherwise (s zero or subnormal, a tie, or the sum leaves the binade)
//    thread 0 adds the chunk's terms one by one from LDS, as the reference.
// Terms are norm_sq in f32 (re*re + im*im without contraction).
constexpr int P0_THREADS = 1024;
constexpr int P0_PER = 16;
constexpr int P0_CHUNK = P0_THREADS * P0_PER;

__device__ __forceinline__ uint64_t ins0(uint64_t k, uint32_t q) {
  return ((k >> q) << (q + 1)) | (k & ((1ull << q) - 1));
}

__global__ __launch_bounds__(P0_THREADS) void k_p0_seq_f32(const float2* __restrict__ a, uint64_t nhalf, uint32_t q,
                                                         float* __restrict__ out) {
  __shared__ __align__(16) float xs[P0_CHUNK];
  __shared__ unsigned long long wsum[P0_THREADS / WAVE];
  __shared__ int wflag[P0_THREADS / WAVE];
  __shared__ float s_sh;
  const int t = threadIdx.x, lane = t & (WAVE - 1), w = t / WAVE;
  if (t == 0) s_sh = 0.0f;
  float2 cur[P0_PER];
  auto load = [&](uint64_t base, float2* v) {
#pragma unroll
    for (int j = 0; j < P0_PER; ++j) {
      uint64_t k = base + (uint64_t)t * P0_PER + j;
      v[j] = k < nhalf ? a[ins0(k, q)] : make_float2(0.0f, 0.0f);
    }
  };
  load(0, cur);
  __syncthreads();
  for (uint64_t base = 0; base < nhalf; base += P0_CHUNK) {
    float x[P0_PER];
#pragma unroll
    for (int j = 0; j < P0_PER; ++j) {
      float re2 = cur[j].x * cur[j].x, im2 = cur[j].y * cur[j].y;
      x[j] = re2 + im2;
    }
    if (base + P0_CHUNK < nhalf) load(base + P0_CHUNK, cur);  // next chunk in flight
#pragma unroll
    for (int j = 0; j < P0_PER; ++j) xs[t * P0_PER + j] = x[j];
    const float s = s_sh;
    const bool normal = s >= 1.17549435e-38f;  // FLT_MIN
    int e2 = 0;
    (void)frexpf(s, &e2);  // s = m 2^e2, m in [0.5, 1): binade E = e2 - 1
    const int E = e2 - 1;
    unsigned long long rs = 0;
    int bad = 0;
    if (normal) {
#pragma unroll
      for (int j = 0; j < P0_PER; ++j) {
        double y = ldexp((double)x[j], 23 - E);  // x / u, exact
        if (!(y < 33554432.0)) {                 // >= 2^25 (or NaN): cannot stay in the binade
          bad = 1;
          continue;
        }
        double f = floor(y), d = y - f;
        if (d == 0.5) bad = 1;
        rs += (unsigned long long)f + (d > 0.5 ? 1ull : 0ull);
      }
    }
#pragma unroll
    for (int off = WAVE / 2; off > 0; off >>= 1) rs += __shfl_down(rs, off, WAVE);
    bad = __any(bad);
    if (lane == 0) {
      wsum[w] = rs;
      wflag[w] = bad;
    }
    __syncthreads();
    if (t == 0) {
      unsigned long long tot = 0;
      int anybad = !normal;
      for (int k = 0; k < P0_THREADS / WAVE; ++k) {
        tot += wsum[k];
        anybad |= wflag[k];
      }
      const unsigned long long s_units = (unsigned long long)ldexpf(s, 23 - E);  // in [2^23, 2^24)
      float ns;
      if (!anybad && s_units + tot < (1ull << 24)) {
        ns = ldexpf((float)(s_units + tot), E - 23);
      } else {
        // terms past the end are +0 and leave the sum unchanged, so whole float4s are read
        ns = s;
        const uint64_t left = nhalf - base;
        const int cnt4 = (int)((left < (uint64_t)P0_CHUNK ? left : (uint64_t)P0_CHUNK) + 3) / 4;
        const float4* x4 = (const float4*)xs;
        for (int j = 0; j < cnt4; ++j) {
          float4 v = x4[j];
          ns = ns + v.x;
          ns = ns + v.y;
          ns = ns + v.z;
          ns = ns + v.w;
        }
      }
      s_sh = ns;
    }
    __syncthreads();
  }
  if (t == 0) *out = s_sh;
}

// probabilities (state.rs:123-125): f64 re*re+im*im; f32 computed in f32 then
// widened (mixed_precision.rs:303-305).
__global__ __launch_bounds__(BLOCK) void k_probs_f64(const double2* __restrict__ a, uint64_t n, double* __restrict__ o) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) {
    double2 x = a[i];
    o[i] = x.x * x.x + x.y * x.y;
  }
}
__global__ __launch_bounds__(BLOCK) void k_probs_f32(const float2* __restrict__ a, uint64_t n, double* __restrict__ o) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) {
    float2 x = a[i];
    o[i] = (double)(x.x * x.x + x.y * x.y);
  }
}

// Sum tree for sampling: node (level k, prefix r) at index 2^k + r holds
// S_k(r) = sum_{i mod 2^k == r} |a_i|^2.  Leaves (level n) are the FP64
// probabilities; level k is the sum of the two halves of level k+1.
template <class V>
__global__ __launch_bounds__(BLOCK) void k_tree_leaves(const V* __restrict__ a, uint64_t n, double* __restrict__ tree) {
  uint64_t i = flat_block() * BLOCK + threadIdx.x;
  if (i < n) tree[n + i] = nsq(a[i]);
}
__global__ __launch_bounds__(BLOCK) void k_tree_level(double* __restrict__ tree, uint64_t width) {
  uint64_t r = flat_block() * BLOCK + threadIdx.x;
  if (r < width) tree[width + r] = tree[2 * width + r] + tree[3 * width + r];
}

// One thread per shot: measure_all (state.rs:329-335) as n conditional
// decisions.  After collapsing qubits 0..k-1 to the bits of r, the reference's
// p0 for qubit k is S_{k+1}(r) divided by the product of the kept
// probabilities (each collapse scales by 1/sqrt(prob)); outcome = draw >= p0,
// prob = outcome ? 1 - p0 : p0 (state.rs:302-304).
__global__ __launch_bounds__(BLOCK) void k_sample(const double* __restrict__ tree, uint32_t nq,
                                                  const double* __restrict__ draws, uint64_t shots,
                                                  uint8_t* __restrict__ bits, unsigned long long* __restrict__ ties) {
  uint64_t s = flat_block() * BLOCK + threadIdx.x;
  if (s >= shots) return;
  uint64_t r = 0;
  double kept = 1.0;
  unsigned long long nt = 0;
  for (uint32_t k = 0; k < nq; ++k) {
    double p0 = tree[(2ull << k) + r] / kept;
    double u = draws[s * nq + k];
    int res = u >= p0;
    if (fabs(u - p0) < 1e-12) ++nt;
    double prob = res ? 1.0 - p0 : p0;
    kept *= prob;
    r |= (uint64_t)res << k;
    bits[s * nq + k] = (uint8_t)res;
  }
  if (nt) atomicAdd(ties, nt);
}

inline unsigned red_blocks(uint64_t n) {
  uint64_t b = (n + BLOCK * 8 - 1) / (BLOCK * 8);
  if (b < 1) b = 1;
  if (b > MAX_RED_BLOCKS) b = MAX_RED_BLOCKS;
  return (unsigned)b;
}

inline dim3 grid_for(uint64_t items) { return grid_blocks((items + BLOCK - 1) / BLOCK); }

// Read `count` doubles from s.red + offset into host `out` (synchronous).
int read_red(State& s, size_t offset, size_t count, double* out) {
  RQ_HIP(hipMemcpyAsync(s.host_red, s.red + offset, count * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  for (size_t k = 0; k < count; ++k) out[k] = s.host_red[k];
  return RQ_OK;
}

// Fixed device scratch for the allocation-free Hamiltonian expectation: the
// per-block partials of one launch of k_pauli (up to MAX_TERMS_PER_LAUNCH
// terms x MAX_RED_BLOCKS blocks) and the folded per-term values.  Each GPU
// holds its own copy for the life of the process; every call on a GPU shares
// it, so calls that overlap on one GPU race.  Every slot a call reads was
// written earlier in the same call, so nothing needs clearing.
constexpr size_t NOALLOC_DOUBLES = (size_t)MAX_TERMS_PER_LAUNCH * (MAX_RED_BLOCKS + 1);
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

// Device address of `scratch` on each GPU, looked up once per device with that
// device current (the host-side symbol is only a stand-in for it).
constexpr int NOALLOC_MAX_DEVICES = 64;
double* g_scratch_ptr[NOALLOC_MAX_DEVICES] = {};
std::mutex g_scratch_mu;

int scratch_for_current_device(double** out) {
  int dev = -1;
  RQ_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fail(RQ_ERR_DEVICE, "device index out of range for the fixed scratch");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) return hip_fail(e == hipSuccess ? hipErrorInvalidSymbol : e, "scratch symbol address");
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return RQ_OK;
}

}  // namespace

int reduce_masked_normsq(State& s, uint64_t mask, uint64_t want, double* out) {
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, nb + 1);
  if (rc) return rc;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_masked_normsq<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n,
                       mask, want, s.red + 1);
  else
    hipLaunchKernelGGL(k_masked_normsq<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, mask,
                       want, s.red + 1);
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(BLOCK), 0, s.stream, s.red + 1, (int)nb, s.red);
  RQ_HIP(hipGetLastError());
  return read_red(s, 0, 1, out);
}

int reduce_p0_seq_f32(State& s, uint32_t q, float* out) {
  int rc = ensure_red(s, 1);
  if (rc) return rc;
  hipLaunchKernelGGL(k_p0_seq_f32, dim3(1), dim3(P0_THREADS), 0, s.stream, (const float2*)s.amps, (uint64_t)(s.n / 2),
                     q, (float*)s.red);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipMemcpyAsync(s.host_red, s.red, sizeof(float), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  memcpy(out, s.host_red, sizeof(float));
  return RQ_OK;
}

int reduce_pauli(State& s, const PauliCanon* terms, size_t nterms, double* out) {
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, (size_t)MAX_TERMS_PER_LAUNCH * (nb + 1));
  if (rc) return rc;
  for (size_t t0 = 0; t0 < nterms; t0 += MAX_TERMS_PER_LAUNCH) {
    int cnt = (int)((nterms - t0) < (size_t)MAX_TERMS_PER_LAUNCH ? (nterms - t0) : MAX_TERMS_PER_LAUNCH);
    PauliBatch b{};
    for (int k = 0; k < cnt; ++k) b.t[k] = PauliDev{terms[t0 + k].x, terms[t0 + k].zy, terms[t0 + k].ny, 0};
    double* part = s.red + MAX_TERMS_PER_LAUNCH;
    if (s.prec == RQ_C64)
      hipLaunchKernelGGL(k_pauli<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, b, cnt,
                         part);
    else
      hipLaunchKernelGGL(k_pauli<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, b, cnt,
                         part);
    hipLaunchKernelGGL(k_fold, dim3(cnt), dim3(BLOCK), 0, s.stream, part, (int)nb, s.red);
    RQ_HIP(hipGetLastError());
    rc = read_red(s, 0, cnt, out + t0);
    if (rc) return rc;
  }
  return RQ_OK;
}

int reduce_hamiltonian_noalloc(State& s, const double* coeffs, const PauliCanon* terms, size_t nterms, double* acc) {
  const unsigned nb = red_blocks(s.n);
  if ((size_t)MAX_TERMS_PER_LAUNCH * (nb + 1) > NOALLOC_DOUBLES)  // nb <= MAX_RED_BLOCKS: never taken
    return fail(RQ_ERR_ARGUMENT, "state too large for the fixed scratch");
  double* base = nullptr;
  int rc = scratch_for_current_device(&base);
  if (rc) return rc;
  const int cnt = (int)nterms;
  if (cnt < 1 || cnt > MAX_TERMS_PER_LAUNCH) return fail(RQ_ERR_ARGUMENT, "term batch out of range");
  PauliBatch b{};
  for (int k = 0; k < cnt; ++k) b.t[k] = PauliDev{terms[k].x, terms[k].zy, terms[k].ny, 0};
  double* part = base + MAX_TERMS_PER_LAUNCH;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_pauli<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, b, cnt,
                       part);
  else
    hipLaunchKernelGGL(k_pauli<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, b, cnt,
                       part);
  hipLaunchKernelGGL(k_fold, dim3(cnt), dim3(BLOCK), 0, s.stream, part, (int)nb, base);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipMemcpyAsync(s.host_red, base, cnt * sizeof(double), hipMemcpyDeviceToHost, s.stream));
  RQ_HIP(hipStreamSynchronize(s.stream));
  for (int k = 0; k < cnt; ++k) *acc += coeffs[k] * s.host_red[k];  // state.rs:403-408, in term order
  return RQ_OK;
}

int reduce_z_all(State& s, double* out) {
  unsigned nb = red_blocks(s.n);
  size_t slots = (size_t)s.nq + 1;
  int rc = ensure_red(s, slots * (nb + 1));
  if (rc) return rc;
  double* part = s.red + slots;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_z_all<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, s.nq, part);
  else
    hipLaunchKernelGGL(k_z_all<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, s.nq, part);
  hipLaunchKernelGGL(k_fold, dim3((unsigned)slots), dim3(BLOCK), 0, s.stream, part, (int)nb, s.red);
  RQ_HIP(hipGetLastError());
  std::vector<double> v(slots);
  rc = read_red(s, 0, slots, v.data());
  if (rc) return rc;
  for (uint32_t q = 0; q < s.nq; ++q) out[q] = v[0] - 2.0 * v[q + 1];
  return RQ_OK;
}

int reduce_inner(const State& a, const State& b, double* re, double* im) {
  State& s = const_cast<State&>(a);
  unsigned nb = red_blocks(s.n);
  int rc = ensure_red(s, 2 * (nb + 1));
  if (rc) return rc;
  // b's stream may differ: order b's pending work before a's reduction.
  if (b.stream != s.stream) RQ_HIP(hipStreamSynchronize(b.stream));
  double* part = s.red + 2;
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_inner<double2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const double2*)a.amps,
                       (const double2*)b.amps, s.n, part);
  else
    hipLaunchKernelGGL(k_inner<float2>, dim3(nb), dim3(BLOCK), 0, s.stream, (const float2*)a.amps,
                       (const float2*)b.amps, s.n, part);
  hipLaunchKernelGGL(k_fold, dim3(2), dim3(BLOCK), 0, s.stream, part, (int)nb, s.red);
  RQ_HIP(hipGetLastError());
  double v[2];
  rc = read_red(s, 0, 2, v);
  if (rc) return rc;
  *re = v[0];
  *im = v[1];
  return RQ_OK;
}

int launch_probabilities(State& s, double* d_out) {
  if (s.prec == RQ_C64)
    hipLaunchKernelGGL(k_probs_f64, grid_for(s.n), dim3(BLOCK), 0, s.stream, (const double2*)s.amps, s.n, d_out);
  else
    hipLaunchKernelGGL(k_probs_f32, grid_for(s.n), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n, d_out);
  RQ_HIP(hipGetLastError());
  return RQ_OK;
}

// The sampler's device memory (sum tree, draws, bits, tie counter) is one
// block.  Blocks up to 4 GiB stay with the state for its next sampling
// (repeated sampling of a 26-qubit state then allocates and frees nothing);
// larger ones are released after the call.
constexpr size_t SAMPLER_KEEP_BYTES = size_t(4) << 30;

int sample_shots(State& s, const double* draws, size_t shots, uint8_t* bits, uint64_t* near_ties) {
  if (shots == 0) {
    if (near_ties) *near_ties = 0;
    return RQ_OK;
  }
  size_t nd = shots * s.nq;
  auto al = [](size_t x) { return (x + 255) & ~(size_t)255; };
  const size_t o_draws = al(2 * s.n * sizeof(double)), o_bits = o_draws + al(nd * sizeof(double)),
               o_ties = o_bits + al(nd), need = o_ties + al(sizeof(unsigned long long));
  hipError_t e = hipSuccess;
  if (s.samp_cap < need) {
    if (s.samp) {
      e = hipStreamSynchronize(s.stream);
      if (e == hipSuccess) e = hipFree(s.samp);
      s.samp = nullptr;
      s.samp_cap = 0;
    }
    if (e == hipSuccess) e = hipMalloc(&s.samp, need);
    if (e == hipSuccess) s.samp_cap = need;
    else s.samp = nullptr;
  }
  unsigned char* base = reinterpret_cast<unsigned char*>(s.samp);
  double* tree = reinterpret_cast<double*>(base);
  double* d_draws = reinterpret_cast<double*>(base + o_draws);
  uint8_t* d_bits = base + o_bits;
  unsigned long long* d_ties = reinterpret_cast<unsigned long long*>(base + o_ties);
  int rc = RQ_OK;
  if (e != hipSuccess) {
    rc = fail(RQ_ERR_ALLOC, "sample_shots: device allocation failed");
  } else {
    e = hipMemcpyAsync(d_draws, draws, nd * sizeof(double), hipMemcpyHostToDevice, s.stream);
    if (e == hipSuccess) e = hipMemsetAsync(d_ties, 0, sizeof(unsigned long long), s.stream);
    if (e == hipSuccess) {
      if (s.prec == RQ_C64)
        hipLaunchKernelGGL(k_tree_leaves<double2>, grid_for(s.n), dim3(BLOCK), 0, s.stream, (const double2*)s.amps,
                           s.n, tree);
      else
        hipLaunchKernelGGL(k_tree_leaves<float2>, grid_for(s.n), dim3(BLOCK), 0, s.stream, (const float2*)s.amps, s.n,
                           tree);
      for (int k = (int)s.nq - 1; k >= 0; --k) {
        uint64_t w = 1ull << k;
        hipLaunchKernelGGL(k_tree_level, grid_for(w), dim3(BLOCK), 0, s.stream, tree, w);
      }
      hipLaunchKernelGGL(k_sample, grid_for(shots), dim3(BLOCK), 0, s.stream, tree, s.nq, d_draws, (uint64_t)shots,
                         d_bits, d_ties);
      e = hipGetLastError();
    }
    unsigned long long ties = 0;
    if (e == hipSuccess) e = hipMemcpyAsync(bits, d_bits, nd, hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipMemcpyAsync(&ties, d_ties, sizeof(ties), hipMemcpyDeviceToHost, s.stream);
    if (e == hipSuccess) e = hipStreamSynchronize(s.stream);
    if (e != hipSuccess) rc = hip_fail(e, "sample_shots");
    else if (near_ties) *near_ties = ties;
  }
  (void)hipStreamSynchronize(s.stream);
  if (s.samp && (rc != RQ_OK || s.samp_cap > SAMPLER_KEEP_BYTES)) {
    (void)hipFree(s.samp);
    s.samp = nullptr;
    s.samp_cap = 0;
  }
  return rc;
}

}  // namespace rq
