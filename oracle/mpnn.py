"""CPU restatement of the learned preconditioner: test infrastructure only.

Only tests/ may import this module. It restates, in numpy, the network and
the training loss of bfm_amd/csrc/hip/mpnn.hip (no reference counterpart:
libbfm has no learned part), in two modes:

  * exact (bf=False): float64 everywhere. Its gradient is checked against
    central finite differences of its own loss, and the operator against
    symmetry, positive definiteness and linearity
    (tests/test_mpnn_oracle.py), which pins the restatement;
  * bf16 (bf=True): the GPU's roundings at the same points (layer weights,
    stored activations and messages, the adjoint's staged inputs and outputs,
    and the backward passes' layer-output gradients are rounded to BF16 with
    round-to-nearest-even; products of BF16 values are summed in at least
    FP32). The GPU's forward and gradient are checked against this mode.

Preconditioner:  z = D^-1 r + P Delta^-1/2 N^T N Delta^-1/2 P r
  N (linear message passing, H = 64):
    h0 = x W0;  m_i = sum_j a_ij[g(c)] h_j[c];  h' = [h, m] W_l  (3 layers)
    a_ij = A_ij / sqrt(delta_i delta_j), channel group g = c // 16 <-> block
    entry (g // 2, g % 2)
  N^T (adjoint): a = v; [s, t] = a W_l^T; a = s + agg^T(t); u = a W0^T
Loss of a batch: mean over pairs of (z - e)^T A (z - e) / (e^T A e), A the
FP32 matrix; the gradient of u = N^T N x is the weight gradient of
<N g, N x> + <N x, N g> with g = P Delta^-1/2 dL/dz.
"""

import numpy as np
import scipy.sparse as sp

H = 64
OFF_W0 = 0
LAYER = 2 * H * H
OFF_L = OFF_W0 + 2 * H
N_PARAMS = OFF_L + 3 * LAYER


def bf16(x):
    """Round to BF16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)
    nan = np.isnan(a)
    if nan.any():
        out = out.copy()
        out[nan] = np.nan
    return out


def unpack(p):
    p = np.asarray(p, np.float64)
    d = {"W0": p[OFF_W0:OFF_L].reshape(2, H)}
    for l in range(3):
        o = OFF_L + l * LAYER
        d[f"W{l + 1}"] = p[o:o + LAYER].reshape(2 * H, H)
    return d


def pack(d):
    p = np.zeros(N_PARAMS)
    p[OFF_W0:OFF_L] = d["W0"].reshape(-1)
    for l in range(3):
        o = OFF_L + l * LAYER
        p[o:o + LAYER] = d[f"W{l + 1}"].reshape(-1)
    return p


def features(rowptr, col, val, fixed):
    """Node features as the GPU computes them (FP64, stored FP32): columns
    dinv00, dinv01, dinv10, dinv11, delta, 1/sqrt(delta), c, fixed bits.
    rowptr/col/val: block-CSR (val (nnzb, 2, 2)) with sorted rows;
    fixed (n, 2) bool."""
    n = len(rowptr) - 1
    rows = np.repeat(np.arange(n), np.diff(rowptr))
    diag = val[rows == col]
    det = diag[:, 0, 0] * diag[:, 1, 1] - diag[:, 0, 1] * diag[:, 1, 0]
    delta = 0.5 * (diag[:, 0, 0] + diag[:, 1, 1])
    f = np.zeros((n, 8))
    f[:, 0] = diag[:, 1, 1] / det
    f[:, 1] = -diag[:, 0, 1] / det
    f[:, 2] = -diag[:, 1, 0] / det
    f[:, 3] = diag[:, 0, 0] / det
    f[:, 4] = delta
    f[:, 5] = 1.0 / np.sqrt(delta)
    ld = np.log(delta.astype(np.float32).astype(np.float64))
    sums = np.add.reduceat(ld[col], rowptr[:-1])
    f[:, 6] = ld - sums / np.diff(rowptr)
    f[:, 7] = fixed[:, 0].astype(float) + 2 * fixed[:, 1].astype(float)
    return f.astype(np.float32)


class Graph:
    """The node graph with the FP32 matrix blocks (block-CSR, internal order)."""

    def __init__(self, rowptr, col, val, feat):
        self.n = len(rowptr) - 1
        self.rowptr, self.col = rowptr, col
        self.val32 = np.asarray(val, np.float32).astype(np.float64)   # (nnzb, 2, 2)
        self.feat = np.asarray(feat, np.float32).astype(np.float64)
        rows = np.repeat(np.arange(self.n), np.diff(rowptr))
        isd = self.feat[:, 5]
        self.groups = []
        for g in range(4):
            a, b = divmod(g, 2)
            # float32 products, in the GPU's order
            w = (self.val32[:, a, b].astype(np.float32) * isd[rows].astype(np.float32)
                 * isd[col].astype(np.float32)).astype(np.float64)
            self.groups.append(sp.csr_matrix((w, (rows, col)), shape=(self.n, self.n)))
        self.A = sp.bsr_matrix((self.val32.copy(), col, rowptr), shape=(2 * self.n, 2 * self.n)).tocsr()
        fx = self.feat[:, 7].astype(int)
        self.free = np.stack([(fx & 1) == 0, (fx & 2) == 0], -1)   # (n, 2)

    def block_jacobi(self, r):
        f = self.feat
        return np.stack([f[:, 0] * r[..., 0] + f[:, 1] * r[..., 1], f[:, 2] * r[..., 0] + f[:, 3] * r[..., 1]], -1)

    def scale(self, v):
        """P Delta^-1/2 v (v (..., n, 2))."""
        return np.where(self.free, self.feat[:, 5][:, None] * v, 0.0)


def _agg(G, h):
    m = np.empty_like(h)
    for g in range(4):
        sl = slice(16 * g, 16 * g + 16)
        for b in range(h.shape[0]):
            m[b, :, sl] = G.groups[g] @ h[b, :, sl]
    return m


def _agg_t(G, t):
    out = np.zeros_like(t)
    for g in range(4):
        sl = slice(16 * g, 16 * g + 16)
        for b in range(t.shape[0]):
            out[b, :, sl] = G.groups[g].T @ t[b, :, sl]
    return out


def _rnd(bf):
    return (lambda v: bf16(v).astype(np.float64)) if bf else (lambda v: v)


def net_forward(P, G, x, bf, act_bf=None):
    """N x for x (B, n, 2): activations h0..h3 and the layer inputs [h, m].
    bf: BF16 layer weights; act_bf (default bf): BF16 activations."""
    wr = _rnd(bf)
    rnd = _rnd(bf if act_bf is None else act_bf)
    h = [rnd(x @ P["W0"])]
    Xs = []
    for l in range(3):
        m = rnd(_agg(G, h[-1]))
        X = np.concatenate([h[-1], m], axis=-1)
        h.append(rnd(X @ wr(P[f"W{l + 1}"])))
        Xs.append(X)
    return h, Xs


def net_adjoint(P, G, v, bf, act_bf=None):
    """N^T v for v (B, n, 64): (B, n, 2)."""
    wr = _rnd(bf)
    rnd = _rnd(bf if act_bf is None else act_bf)
    a = v
    s = t = None
    for l in (2, 1, 0):
        if l < 2:
            a = rnd(s + _agg_t(G, t))
        O = a @ wr(P[f"W{l + 1}"]).T
        s, t = rnd(O[..., :H]), rnd(O[..., H:])
    a0 = s + _agg_t(G, t)
    return a0 @ P["W0"].T


def apply(params, G, r, bf=True, act_bf=False):
    """z = D^-1 r + P Delta^-1/2 N^T N Delta^-1/2 P r for r (n, 2) or (B, n, 2).
    As in a solve: BF16 layer weights (bf), exact activations (act_bf False;
    the GPU keeps them in FP32 with split MFMA inputs)."""
    P = unpack(params)
    r = np.asarray(r, np.float64)
    one = r.ndim == 2
    rb = r[None] if one else r
    x = G.scale(rb)
    h, _ = net_forward(P, G, x, bf, act_bf)
    u = net_adjoint(P, G, h[3], bf, act_bf)
    z = G.block_jacobi(rb) + G.scale(u)
    return z[0] if one else z


def _weight_grad(P, G, Xs, x, up, bf):
    rnd = _rnd(bf)
    grad = {}
    dh = up
    for l in (2, 1, 0):
        dpre = rnd(dh)
        grad[f"W{l + 1}"] = np.einsum("bik,bio->ko", Xs[l], dpre)
        dX = dpre @ rnd(P[f"W{l + 1}"]).T
        dh = dX[..., :H] + _agg_t(G, dX[..., H:])
    grad["W0"] = np.einsum("bia,bio->ao", x, dh)
    return pack(grad)


def loss_grad(params, G, r, e, den, bf=True):
    """Mean A-norm loss over the batch (r, e (B, n, 2), den (B,)) and its gradient."""
    P = unpack(params)
    r = np.asarray(r, np.float64)
    e = np.asarray(e, np.float64)
    den = np.asarray(den, np.float64)
    B = r.shape[0]
    x = G.scale(r)
    if bf:
        x = x.astype(np.float32).astype(np.float64)
    h, Xs = net_forward(P, G, x, bf)
    u = net_adjoint(P, G, h[3], bf)
    z = G.block_jacobi(r) + G.scale(u)
    d = z - e
    loss = 0.0
    dz = np.zeros_like(z)
    for b in range(B):
        q = (G.A @ d[b].reshape(-1)).reshape(-1, 2)
        loss += (d[b] * q).sum() / (den[b] * B)
        dz[b] = 2.0 * q / (den[b] * B)
    g = G.scale(dz)
    if bf:
        g = g.astype(np.float32).astype(np.float64)
    hg, Xg = net_forward(P, G, g, bf)
    grad = _weight_grad(P, G, Xs, x, hg[3], bf) + _weight_grad(P, G, Xg, g, h[3], bf)
    return loss, grad


def init_params(seed=0, encoder_scale=0.0):
    """Glorot-uniform layers (a different generator from the GPU's
    model_create: tests pass the GPU's parameters explicitly); the encoder is
    encoder_scale times a standard normal draw (0: block Jacobi)."""
    rng = np.random.default_rng(seed)
    a = np.sqrt(6.0 / (3 * H))
    d = {"W0": encoder_scale * rng.standard_normal((2, H))}
    for l in range(3):
        d[f"W{l + 1}"] = rng.uniform(-a, a, (2 * H, H))
    return pack(d)
