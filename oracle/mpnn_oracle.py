"""FP64 numpy restatement of the warm-start graph network (test infrastructure).

The reference has no learned component; this oracle is the network's own
specification, pinned by finite differences (tests/test_mpnn_oracle.py), and
the device kernels (ruqu_amd/csrc/kernels_mpnn.hip) are checked against it.
With bf16=True it rounds to bfloat16 (round to nearest even) at exactly the
points where the device feeds a matrix-core product, so forward values and
gradients can be compared at FP32-accumulation tolerance.

Network (hidden H = 256, 3 message-passing layers, depths p = 1..8):
  node features x_v = [deg_v / 16, wdeg_v / 8, 0...] (KIN = 32 columns)
  h0 = relu(x W_in + b_in)
  layer k = 1..3:  m = A h^{k-1}  (A_vu = w_uv, undirected)
                   Z_k = [h^{k-1} | m] W_k + b_k       (W_k: 2H x H)
                   h^k = h^{k-1} + relu(Z_k)
  g = mean over the graph's nodes of h^3
  y = [g | onehot(p) | n/22 | edge density | 0...]     (HIN = H + 32)
  U = relu(y W_h1 + b_h1);  O = U W_h2 + b_h2           (W_h2: H x 64)
  gammas = O[:, 0:8], betas = O[:, 8:16]; entries l >= p are unused.
  loss = sum over used entries (O - target)^2 / (2 * sum_g p_g)
"""
import numpy as np

H = 256
KIN = 32
GF = 32
HIN = H + GF
NOUT = 64
PMAX = 8
LAYERS = 3

# Flat parameter layout (FP32 on the device): name, shape.
LAYOUT = [("W_in", (KIN, H)), ("b_in", (H,))]
for _k in range(LAYERS):
    LAYOUT += [("W_%d" % (_k + 1), (2 * H, H)), ("b_%d" % (_k + 1), (H,))]
LAYOUT += [("W_h1", (HIN, H)), ("b_h1", (H,)), ("W_h2", (H, NOUT)), ("b_h2", (NOUT,))]


def offsets():
    out, o = {}, 0
    for name, shape in LAYOUT:
        out[name] = (o, shape)
        o += int(np.prod(shape))
    return out, o


OFFS, NPARAMS = offsets()


def unpack(flat):
    return {k: flat[o:o + int(np.prod(s))].reshape(s) for k, (o, s) in OFFS.items()}


def bf16(x):
    """Round float32 values to bfloat16 (nearest even), returned as float64."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return (r.astype(np.uint32).view(np.float32)).astype(np.float64)


class Batch:
    """Graphs packed for the network: nodes of graph g are node_off[g] ..
    node_off[g+1]-1; each undirected edge appears as (u->v) and (v->u)."""

    def __init__(self, graphs, ps):
        self.ngraphs = len(graphs)
        self.p = np.asarray(ps, dtype=np.int32)
        off = [0]
        src, dst, w, feat, gfeat = [], [], [], [], []
        for (n, edges, weights), p in zip(graphs, ps):
            base = off[-1]
            deg = np.zeros(n)
            wdeg = np.zeros(n)
            for (i, j), x in zip(edges, weights):
                src += [base + i, base + j]
                dst += [base + j, base + i]
                w += [x, x]
                deg[i] += 1
                deg[j] += 1
                wdeg[i] += x
                wdeg[j] += x
            f = np.zeros((n, KIN))
            f[:, 0] = deg / 16.0
            f[:, 1] = wdeg / 8.0
            feat.append(f)
            gf = np.zeros(GF)
            gf[p - 1] = 1.0
            gf[PMAX] = n / 22.0
            gf[PMAX + 1] = len(edges) / (n * (n - 1) / 2.0)
            gfeat.append(gf)
            off.append(base + n)
        self.node_off = np.asarray(off, dtype=np.int32)
        self.nnodes = off[-1]
        src, dst, w = np.asarray(src, dtype=np.int64), np.asarray(dst, dtype=np.int64), np.asarray(w)
        order = np.lexsort((src, dst))  # CSR by destination, sources ascending
        self.src, self.dst, self.w = src[order], dst[order], w[order]
        self.row_ptr = np.zeros(self.nnodes + 1, dtype=np.int32)
        np.add.at(self.row_ptr, self.dst + 1, 1)
        self.row_ptr = np.cumsum(self.row_ptr).astype(np.int32)
        self.x = np.concatenate(feat).astype(np.float32)
        self.gfeat = np.asarray(gfeat, dtype=np.float32)
        self.graph_of = np.repeat(np.arange(self.ngraphs), np.diff(self.node_off))

    def aggregate(self, h):
        """m_v = sum_{u->v} w_uv h_u (the operator is symmetric)."""
        m = np.zeros_like(h)
        np.add.at(m, self.dst, self.w[:, None] * h[self.src])
        return m

    def mean_pool(self, h):
        g = np.zeros((self.ngraphs, h.shape[1]))
        np.add.at(g, self.graph_of, h)
        return g / np.diff(self.node_off)[:, None]


def _r(x, on):
    return bf16(x) if on else np.asarray(x, dtype=np.float64)


def forward(flat, b, bf=False, keep=False):
    P = unpack(np.asarray(flat, dtype=np.float64))
    c = {}
    x = np.asarray(b.x, dtype=np.float64)
    Wr = {k: _r(P[k], bf) for k in P if k.startswith("W")}
    z0 = _r(x, bf) @ Wr["W_in"] + P["b_in"]
    h = np.maximum(z0, 0.0)
    c["z0"] = z0
    hs, xs, zs = [h], [], []
    for k in range(LAYERS):
        if bf:
            h = h.astype(np.float32).astype(np.float64)  # FP32 state on the device
        m = b.aggregate(h)
        X = np.concatenate([_r(h, bf), _r(m, bf)], axis=1)
        Z = X @ Wr["W_%d" % (k + 1)] + P["b_%d" % (k + 1)]
        h = h + np.maximum(Z, 0.0)
        xs.append(X)
        zs.append(Z)
        hs.append(h)
    g = b.mean_pool(h)
    y = np.concatenate([g, np.asarray(b.gfeat, dtype=np.float64)], axis=1)
    yr = _r(y, bf)
    zu = yr @ Wr["W_h1"] + P["b_h1"]
    U = np.maximum(zu, 0.0)
    O = _r(U, bf) @ Wr["W_h2"] + P["b_h2"]
    if keep:
        c.update(hs=hs, xs=xs, zs=zs, g=g, yr=yr, zu=zu, U=U, O=O, Wr=Wr)
        return O, c
    return O


def loss_grad_out(O, targets, p):
    """Masked MSE and its gradient with respect to O (targets: [G, 16])."""
    G = O.shape[0]
    mask = np.zeros((G, NOUT))
    for g in range(G):
        mask[g, : p[g]] = 1.0
        mask[g, PMAX: PMAX + p[g]] = 1.0
    T = np.zeros((G, NOUT))
    T[:, : 2 * PMAX] = targets
    denom = 2.0 * float(np.sum(p))
    d = (O - T) * mask
    return float(np.sum(d * d) / denom), 2.0 * d / denom


def backward(flat, b, targets, bf=False):
    """Loss and d loss / d params (flat, FP64), backpropagated by hand."""
    O, c = forward(flat, b, bf=bf, keep=True)
    loss, dO = loss_grad_out(O, targets, b.p)
    P = unpack(np.asarray(flat, dtype=np.float64))
    Wr = c["Wr"]
    grads = {}
    dOr = _r(dO, bf)
    Ur = _r(c["U"], bf)
    grads["W_h2"] = Ur.T @ dOr
    grads["b_h2"] = dO.sum(0)
    dU = dOr @ Wr["W_h2"].T
    dzu = dU * (c["zu"] > 0)
    dzur = _r(dzu, bf)
    grads["W_h1"] = c["yr"].T @ dzur
    grads["b_h1"] = dzu.sum(0)
    dy = dzur @ Wr["W_h1"].T
    dg = dy[:, :H]
    dh = dg[b.graph_of] / np.diff(b.node_off)[b.graph_of][:, None]
    for k in reversed(range(LAYERS)):
        Z, X = c["zs"][k], c["xs"][k]
        dZ = dh * (Z > 0)
        dZr = _r(dZ, bf)
        grads["W_%d" % (k + 1)] = X.T @ dZr
        grads["b_%d" % (k + 1)] = dZ.sum(0)
        dX = dZr @ Wr["W_%d" % (k + 1)].T
        dh = dh + dX[:, :H] + b.aggregate(dX[:, H:])
    dz0 = dh * (c["z0"] > 0)
    dz0r = _r(dz0, bf)
    grads["W_in"] = _r(np.asarray(b.x, dtype=np.float64), bf).T @ dz0r
    grads["b_in"] = dz0.sum(0)
    flat_g = np.zeros(NPARAMS)
    for k, (o, s) in OFFS.items():
        flat_g[o:o + int(np.prod(s))] = np.asarray(grads[k]).reshape(-1)
    return loss, flat_g


def init_params(seed):
    """He-normal weights from numpy's PCG64(seed), zero biases (b_h2 = 0.5:
    angles start near the plain optimizer's initial 0.5)."""
    rng = np.random.default_rng(seed)
    flat = np.zeros(NPARAMS)
    for name, (o, s) in OFFS.items():
        if name.startswith("W"):
            fan_in = s[0]
            flat[o:o + int(np.prod(s))] = rng.normal(0.0, np.sqrt(2.0 / fan_in), int(np.prod(s)))
    o, s = OFFS["W_h2"]
    flat[o:o + int(np.prod(s))] *= 0.1
    o, s = OFFS["b_h2"]
    flat[o:o + PMAX * 2] = 0.5
    return flat.astype(np.float32)
