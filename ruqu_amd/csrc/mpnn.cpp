// Warm-start graph network: host driver of the forward pass, the hand-written
// backward pass and Adam (kernels in kernels_mpnn.hip).  No reference
// counterpart; the specification it follows is oracle/mpnn_oracle.py.
//
// Per batch of graphs (N nodes, G graphs):
//   z0 = xin W_in + b_in, h0 = relu(z0)
//   layer k: m = A h, X_k = [h | m] (bf16), Z_k = X_k W_k + b_k, h += relu(Z_k)
//   Y = [mean-pool(h) | graph features] (bf16), zu = Y W_h1 + b_h1,
//   O = bf16(relu(zu)) W_h2 + b_h2
// The backward pass mirrors it (aggregation is its own transpose: the edge
// list holds both directions with the same weight).
#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "internal.h"
#include "mpnn.h"

namespace rq {
namespace mpnn {

void gemm(hipStream_t st, bool a_kc, bool b_kc, int M, int N, int K, const uint16_t* A, int lda, const uint16_t* B,
          int ldb, float* C, int ldc, const float* bias);
void aggregate(hipStream_t st, int nnodes, const int32_t* row_ptr, const int32_t* col, const float* w,
               const float* in, int ldi, float* out, int ldo, bool accumulate);
void pack_x(hipStream_t st, int n, const float* h, const float* m, uint16_t* X);
void relu_add(hipStream_t st, size_t n, const float* z, const float* h_in, float* h_out);
void pool_head_in(hipStream_t st, int G, const int32_t* node_off, const float* h, const float* gfeat, uint16_t* Y);
void relu_bf16(hipStream_t st, size_t n, const float* z, uint16_t* out);
void to_bf16(hipStream_t st, size_t n, const float* x, uint16_t* out);
void loss(hipStream_t st, int G, const float* O, const float* T, const int32_t* p, float inv_denom, float* dO,
          uint16_t* dOr, float* part);
void relu_grad(hipStream_t st, int rows, int width, const float* up, int ldu, const float* z, float* d, uint16_t* dbf);
void unpool(hipStream_t st, int G, const int32_t* node_off, const float* dY, int ldy, float* dh);
void add_left(hipStream_t st, int nnodes, const float* dX, float* dh);
void colsum(hipStream_t st, int rows, int cols, const float* x, int ld, float* out);
void adam(hipStream_t st, size_t n, float* prm, const float* g, float* m1, float* m2, float lr, float b1, float b2,
          float eps, float c1, float c2, uint16_t* pbf);

namespace {

struct Buf {
  void* p = nullptr;
  size_t cap = 0;
};

}  // namespace

struct Net {
  hipStream_t st = nullptr;
  float *prm = nullptr, *grad = nullptr, *m1 = nullptr, *m2 = nullptr;
  uint16_t* pbf = nullptr;
  int steps = 0;
  std::vector<Buf> bufs;
  std::vector<float> host_part;
  // batch views into the workspace (set by upload)
  int G = 0, N = 0;
  int64_t E = 0;
  int32_t *node_off, *row_ptr, *col, *p;
  float *w, *xin, *gfeat, *targets;
  uint16_t* xin_bf;
  float *z0, *h[LAYERS + 1], *m, *Z[LAYERS], *zu, *O;
  uint16_t *X[LAYERS], *Y, *Ur;
  float *dO, *part, *dU, *dzu, *dY, *dh, *dZ, *dX, *dz0;
  uint16_t *dOr, *dzu_bf, *dZ_bf, *dz0_bf;
};

namespace {

template <class T>
int take(Net& n, size_t idx, T** out, size_t count) {
  if (n.bufs.size() <= idx) n.bufs.resize(idx + 1);
  Buf& b = n.bufs[idx];
  size_t cap_bytes = b.cap;
  const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  if (bytes > cap_bytes || !b.p) {
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    RQ_HIP(hipMalloc(&b.p, bytes));
    b.cap = bytes;
  }
  *out = (T*)b.p;
  return RQ_OK;
}

int upload_batch(Net& n, const rq_mpnn_batch* bt, const float* targets) {
  if (!bt || bt->ngraphs <= 0 || bt->nnodes <= 0 || !bt->node_off || !bt->row_ptr || !bt->x || !bt->gfeat || !bt->p ||
      (bt->nedges && (!bt->col || !bt->w)))
    return fail(RQ_ERR_ARGUMENT, "incomplete graph batch");
  for (int g = 0; g < bt->ngraphs; ++g) {
    if (bt->node_off[g + 1] <= bt->node_off[g]) return fail(RQ_ERR_ARGUMENT, "graph without nodes");
    if (bt->p[g] < 1 || bt->p[g] > PMAX) return fail(RQ_ERR_ARGUMENT, "depth p outside 1..8");
  }
  if (bt->node_off[bt->ngraphs] != bt->nnodes || bt->row_ptr[bt->nnodes] != bt->nedges)
    return fail(RQ_ERR_ARGUMENT, "node_off / row_ptr totals do not match nnodes / nedges");
  for (int64_t e = 0; e < bt->nedges; ++e)
    if (bt->col[e] < 0 || bt->col[e] >= bt->nnodes) return fail(RQ_ERR_ARGUMENT, "edge source out of range");
  const int G = bt->ngraphs, N = bt->nnodes;
  const int64_t E = bt->nedges;
  n.G = G;
  n.N = N;
  n.E = E;
  size_t i = 0;
  int rc = 0;
  rc |= take(n, i++, &n.node_off, G + 1);
  rc |= take(n, i++, &n.row_ptr, N + 1);
  rc |= take(n, i++, &n.col, E);
  rc |= take(n, i++, &n.w, E);
  rc |= take(n, i++, &n.p, G);
  rc |= take(n, i++, &n.xin, (size_t)N * KIN);
  rc |= take(n, i++, &n.xin_bf, (size_t)N * KIN);
  rc |= take(n, i++, &n.gfeat, (size_t)G * GF);
  rc |= take(n, i++, &n.targets, (size_t)G * 2 * PMAX);
  rc |= take(n, i++, &n.z0, (size_t)N * H);
  for (int k = 0; k <= LAYERS; ++k) rc |= take(n, i++, &n.h[k], (size_t)N * H);
  rc |= take(n, i++, &n.m, (size_t)N * H);
  for (int k = 0; k < LAYERS; ++k) {
    rc |= take(n, i++, &n.X[k], (size_t)N * 2 * H);
    rc |= take(n, i++, &n.Z[k], (size_t)N * H);
  }
  rc |= take(n, i++, &n.Y, (size_t)G * HIN);
  rc |= take(n, i++, &n.zu, (size_t)G * H);
  rc |= take(n, i++, &n.Ur, (size_t)G * H);
  rc |= take(n, i++, &n.O, (size_t)G * NOUT);
  rc |= take(n, i++, &n.dO, (size_t)G * NOUT);
  rc |= take(n, i++, &n.dOr, (size_t)G * NOUT);
  rc |= take(n, i++, &n.part, G);
  rc |= take(n, i++, &n.dU, (size_t)G * H);
  rc |= take(n, i++, &n.dzu, (size_t)G * H);
  rc |= take(n, i++, &n.dzu_bf, (size_t)G * H);
  rc |= take(n, i++, &n.dY, (size_t)G * H);
  rc |= take(n, i++, &n.dh, (size_t)N * H);
  rc |= take(n, i++, &n.dZ, (size_t)N * H);
  rc |= take(n, i++, &n.dZ_bf, (size_t)N * H);
  rc |= take(n, i++, &n.dX, (size_t)N * 2 * H);
  rc |= take(n, i++, &n.dz0, (size_t)N * H);
  rc |= take(n, i++, &n.dz0_bf, (size_t)N * H);
  if (rc) return rc;
  RQ_HIP(hipMemcpyAsync(n.node_off, bt->node_off, (G + 1) * sizeof(int32_t), hipMemcpyHostToDevice, n.st));
  RQ_HIP(hipMemcpyAsync(n.row_ptr, bt->row_ptr, (N + 1) * sizeof(int32_t), hipMemcpyHostToDevice, n.st));
  if (E) {
    RQ_HIP(hipMemcpyAsync(n.col, bt->col, E * sizeof(int32_t), hipMemcpyHostToDevice, n.st));
    RQ_HIP(hipMemcpyAsync(n.w, bt->w, E * sizeof(float), hipMemcpyHostToDevice, n.st));
  }
  RQ_HIP(hipMemcpyAsync(n.p, bt->p, G * sizeof(int32_t), hipMemcpyHostToDevice, n.st));
  RQ_HIP(hipMemcpyAsync(n.xin, bt->x, (size_t)N * KIN * sizeof(float), hipMemcpyHostToDevice, n.st));
  RQ_HIP(hipMemcpyAsync(n.gfeat, bt->gfeat, (size_t)G * GF * sizeof(float), hipMemcpyHostToDevice, n.st));
  if (targets)
    RQ_HIP(hipMemcpyAsync(n.targets, targets, (size_t)G * 2 * PMAX * sizeof(float), hipMemcpyHostToDevice, n.st));
  to_bf16(n.st, (size_t)N * KIN, n.xin, n.xin_bf);
  return RQ_OK;
}

void forward(Net& n) {
  const int N = n.N, G = n.G;
  gemm(n.st, true, false, N, H, KIN, n.xin_bf, KIN, n.pbf + OFF_W_IN, H, n.z0, H, n.prm + OFF_B_IN);
  relu_add(n.st, (size_t)N * H, n.z0, nullptr, n.h[0]);
  for (int k = 0; k < LAYERS; ++k) {
    aggregate(n.st, N, n.row_ptr, n.col, n.w, n.h[k], H, n.m, H, false);
    pack_x(n.st, N, n.h[k], n.m, n.X[k]);
    gemm(n.st, true, false, N, H, 2 * H, n.X[k], 2 * H, n.pbf + off_w(k), H, n.Z[k], H, n.prm + off_b(k));
    relu_add(n.st, (size_t)N * H, n.Z[k], n.h[k], n.h[k + 1]);
  }
  pool_head_in(n.st, G, n.node_off, n.h[LAYERS], n.gfeat, n.Y);
  gemm(n.st, true, false, G, H, HIN, n.Y, HIN, n.pbf + OFF_W_H1, H, n.zu, H, n.prm + OFF_B_H1);
  relu_bf16(n.st, (size_t)G * H, n.zu, n.Ur);
  gemm(n.st, true, false, G, NOUT, H, n.Ur, H, n.pbf + OFF_W_H2, NOUT, n.O, NOUT, n.prm + OFF_B_H2);
}

// Gradient of the masked MSE into n.grad; per-graph loss partials in n.part.
void backward(Net& n, float inv_denom) {
  const int N = n.N, G = n.G;
  float* gr = n.grad;
  loss(n.st, G, n.O, n.targets, n.p, inv_denom, n.dO, n.dOr, n.part);
  gemm(n.st, false, false, H, NOUT, G, n.Ur, H, n.dOr, NOUT, gr + OFF_W_H2, NOUT, nullptr);
  colsum(n.st, G, NOUT, n.dO, NOUT, gr + OFF_B_H2);
  gemm(n.st, true, true, G, H, NOUT, n.dOr, NOUT, n.pbf + OFF_W_H2, NOUT, n.dU, H, nullptr);
  relu_grad(n.st, G, H, n.dU, H, n.zu, n.dzu, n.dzu_bf);
  gemm(n.st, false, false, HIN, H, G, n.Y, HIN, n.dzu_bf, H, gr + OFF_W_H1, H, nullptr);
  colsum(n.st, G, H, n.dzu, H, gr + OFF_B_H1);
  gemm(n.st, true, true, G, H, H, n.dzu_bf, H, n.pbf + OFF_W_H1, H, n.dY, H, nullptr);
  unpool(n.st, G, n.node_off, n.dY, H, n.dh);
  for (int k = LAYERS - 1; k >= 0; --k) {
    relu_grad(n.st, N, H, n.dh, H, n.Z[k], n.dZ, n.dZ_bf);
    gemm(n.st, false, false, 2 * H, H, N, n.X[k], 2 * H, n.dZ_bf, H, gr + off_w(k), H, nullptr);
    colsum(n.st, N, H, n.dZ, H, gr + off_b(k));
    gemm(n.st, true, true, N, 2 * H, H, n.dZ_bf, H, n.pbf + off_w(k), H, n.dX, 2 * H, nullptr);
    add_left(n.st, N, n.dX, n.dh);
    aggregate(n.st, N, n.row_ptr, n.col, n.w, n.dX + H, 2 * H, n.dh, H, true);
  }
  relu_grad(n.st, N, H, n.dh, H, n.z0, n.dz0, n.dz0_bf);
  gemm(n.st, false, false, KIN, H, N, n.xin_bf, KIN, n.dz0_bf, H, gr + OFF_W_IN, H, nullptr);
  colsum(n.st, N, H, n.dz0, H, gr + OFF_B_IN);
}

float inv_denom_of(const rq_mpnn_batch* bt) {
  double s = 0;
  for (int g = 0; g < bt->ngraphs; ++g) s += bt->p[g];
  return (float)(1.0 / (2.0 * s));
}

int read_loss(Net& n, float* out) {
  n.host_part.resize(n.G);
  RQ_HIP(hipMemcpyAsync(n.host_part.data(), n.part, n.G * sizeof(float), hipMemcpyDeviceToHost, n.st));
  RQ_HIP(hipStreamSynchronize(n.st));
  double s = 0;
  for (float v : n.host_part) s += v;
  if (out) *out = (float)s;
  return RQ_OK;
}

}  // namespace
}  // namespace mpnn
}  // namespace rq

using namespace rq;
using namespace rq::mpnn;

extern "C" {

size_t rq_mpnn_num_params(void) { return NPARAMS; }

int rq_mpnn_create(rq_mpnn** out) {
  if (!out) return fail(RQ_ERR_ARGUMENT, "null pointer");
  Net* n = new Net();
  hipError_t e = hipStreamCreateWithFlags(&n->st, hipStreamNonBlocking);
  if (e == hipSuccess) e = hipMalloc((void**)&n->prm, NPARAMS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&n->grad, NPARAMS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&n->m1, NPARAMS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&n->m2, NPARAMS * sizeof(float));
  if (e == hipSuccess) e = hipMalloc((void**)&n->pbf, NPARAMS * sizeof(uint16_t));
  if (e != hipSuccess) {
    rq_mpnn_destroy((rq_mpnn*)n);
    return fail(RQ_ERR_DEVICE, std::string("mpnn allocation: ") + hipGetErrorString(e));
  }
  *out = (rq_mpnn*)n;
  return RQ_OK;
}

void rq_mpnn_destroy(rq_mpnn* h) {
  Net* n = (Net*)h;
  if (!n) return;
  if (n->st) (void)hipStreamSynchronize(n->st);
  for (void* p : {(void*)n->prm, (void*)n->grad, (void*)n->m1, (void*)n->m2, (void*)n->pbf})
    if (p) (void)hipFree(p);
  for (Buf& b : n->bufs)
    if (b.p) (void)hipFree(b.p);
  if (n->st) (void)hipStreamDestroy(n->st);
  delete n;
}

int rq_mpnn_set_params(rq_mpnn* h, const float* params) {
  Net* n = (Net*)h;
  if (!n || !params) return fail(RQ_ERR_ARGUMENT, "null pointer");
  RQ_HIP(hipMemcpyAsync(n->prm, params, NPARAMS * sizeof(float), hipMemcpyHostToDevice, n->st));
  RQ_HIP(hipMemsetAsync(n->m1, 0, NPARAMS * sizeof(float), n->st));
  RQ_HIP(hipMemsetAsync(n->m2, 0, NPARAMS * sizeof(float), n->st));
  to_bf16(n->st, NPARAMS, n->prm, n->pbf);
  n->steps = 0;
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipStreamSynchronize(n->st));
  return RQ_OK;
}

int rq_mpnn_get_params(const rq_mpnn* h, float* params) {
  const Net* n = (const Net*)h;
  if (!n || !params) return fail(RQ_ERR_ARGUMENT, "null pointer");
  RQ_HIP(hipMemcpyAsync(params, n->prm, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, n->st));
  RQ_HIP(hipStreamSynchronize(n->st));
  return RQ_OK;
}

int rq_mpnn_forward(rq_mpnn* h, const rq_mpnn_batch* bt, float* angles) {
  Net* n = (Net*)h;
  if (!n || !angles) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = upload_batch(*n, bt, nullptr);
  if (rc) return rc;
  forward(*n);
  RQ_HIP(hipGetLastError());
  RQ_HIP(hipMemcpy2DAsync(angles, 2 * PMAX * sizeof(float), n->O, NOUT * sizeof(float), 2 * PMAX * sizeof(float),
                          n->G, hipMemcpyDeviceToHost, n->st));
  RQ_HIP(hipStreamSynchronize(n->st));
  return RQ_OK;
}

int rq_mpnn_gradient(rq_mpnn* h, const rq_mpnn_batch* bt, const float* targets, float* loss_out, float* grad) {
  Net* n = (Net*)h;
  if (!n || !targets) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = upload_batch(*n, bt, targets);
  if (rc) return rc;
  forward(*n);
  backward(*n, inv_denom_of(bt));
  RQ_HIP(hipGetLastError());
  if (grad) RQ_HIP(hipMemcpyAsync(grad, n->grad, NPARAMS * sizeof(float), hipMemcpyDeviceToHost, n->st));
  return read_loss(*n, loss_out);
}

int rq_mpnn_train_step(rq_mpnn* h, const rq_mpnn_batch* bt, const float* targets, float lr, float* loss_out) {
  Net* n = (Net*)h;
  if (!n || !targets) return fail(RQ_ERR_ARGUMENT, "null pointer");
  int rc = upload_batch(*n, bt, targets);
  if (rc) return rc;
  forward(*n);
  backward(*n, inv_denom_of(bt));
  const float b1 = 0.9f, b2 = 0.999f;
  n->steps += 1;
  const float c1 = 1.0f - std::pow(b1, (float)n->steps), c2 = 1.0f - std::pow(b2, (float)n->steps);
  adam(n->st, NPARAMS, n->prm, n->grad, n->m1, n->m2, lr, b1, b2, 1e-8f, c1, c2, n->pbf);
  RQ_HIP(hipGetLastError());
  return read_loss(*n, loss_out);
}

}  // extern "C"
