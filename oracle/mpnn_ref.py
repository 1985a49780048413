"""CPU reference of the learned SSH preconditioner (TEST INFRASTRUCTURE only).

FESOM2 has no learned preconditioner; this module defines the network the
device kernels implement, so the tests can check the device forward pass and
gradients against it. It is never on the product path.

Network (one sample, N nodes, width C = 64):
    D = diag(A),  r~ = D^-1/2 r,  s = rms(r~),  u = r~ / s
    h0 = relu(u * W0[0] + W0[1])                           (FP32 math, bf16 store)
    for l = 1..3:
        m  = Ahat h_{l-1}                                   (FP32 sum, bf16 store)
        h_l = relu([h_{l-1}, m] @ W_l + b_l)                (bf16 x bf16 -> FP32, bf16 store)
    y = h3 @ w_out + b_out                                  (FP32)
    z = D^-1/2 (r~ + s y)
with Ahat_ij = a_ij / sqrt(a_ii a_jj) over the off-diagonal entries of A.
With all weights zero the network is the diagonal (Jacobi) preconditioner.

Loss over a batch of right-hand sides b with exact solutions x = A^-1 b:
    L = mean_b (z - x)^T A (z - x) / (x^T A x),   z = M(b)
Gradients flow in FP32; the products whose operands are bf16 on the device
(W_l, the layer inputs [h, m], the pre-activation gradients) are rounded to
bf16 here at the same points, unless `exact=True` (FP64 everywhere, for the
finite-difference check of the analytic gradient).
"""
from __future__ import annotations

import numpy as np

C = 64
NLAYERS = 3


def bf16(x):
    """Round float32 values to bfloat16 (nearest, ties to even), returned as float32."""
    a = np.ascontiguousarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    out = r.astype(np.uint32).view(np.float32)
    return np.where(np.isnan(a), a, out)


def param_shapes():
    shapes = [("W0", (2, C))]
    for l in range(1, NLAYERS + 1):
        shapes += [(f"W{l}", (2 * C, C)), (f"b{l}", (C,))]
    shapes += [("w_out", (C,)), ("b_out", (1,))]
    return shapes


def n_params():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def unflatten(theta):
    p, o = {}, 0
    for name, s in param_shapes():
        n = int(np.prod(s))
        p[name] = np.asarray(theta[o:o + n]).reshape(s)
        o += n
    return p


def flatten(p):
    return np.concatenate([np.asarray(p[name]).ravel() for name, _ in param_shapes()])


def init_params(seed=0):
    """He-style initialisation of the hidden layers; output weights start at zero
    so the untrained network equals the diagonal preconditioner."""
    rng = np.random.default_rng(seed)
    p = {"W0": (rng.normal(0, 1.0, (2, C))).astype(np.float32)}
    for l in range(1, NLAYERS + 1):
        p[f"W{l}"] = rng.normal(0, np.sqrt(2.0 / (2 * C)), (2 * C, C)).astype(np.float32)
        p[f"b{l}"] = np.zeros(C, np.float32)
    p["w_out"] = np.zeros(C, np.float32)
    p["b_out"] = np.zeros(1, np.float32)
    return p


class Graph:
    """Off-diagonal CSR of Ahat plus the diagonal of A (CSR of A, diagonal first)."""

    def __init__(self, rowptr, colind, values):
        n = len(rowptr) - 1
        self.n = n
        self.diag = values[rowptr[:-1]].astype(np.float64)
        self.isq = 1.0 / np.sqrt(self.diag)
        rows = np.repeat(np.arange(n), np.diff(rowptr))
        off = colind != rows
        self.rows = rows[off]
        self.cols = colind[off]
        self.w = (values[off] * self.isq[self.rows] * self.isq[self.cols]).astype(np.float32)
        self.rowptr, self.colind, self.values = rowptr, colind, values

    def agg(self, h, dtype=np.float32):
        """sum_j Ahat_ij h_j, rows in CSR order."""
        out = np.zeros((self.n, h.shape[1]), dtype)
        np.add.at(out, self.rows, (self.w[:, None].astype(dtype) * h[self.cols].astype(dtype)))
        return out

    def agg_t(self, h, dtype=np.float32):
        """sum_i Ahat_ij h_i (the transpose, used by the backward pass)."""
        out = np.zeros((self.n, h.shape[1]), dtype)
        np.add.at(out, self.cols, (self.w[:, None].astype(dtype) * h[self.rows].astype(dtype)))
        return out

    def spmv(self, x):
        rows = np.repeat(np.arange(self.n), np.diff(self.rowptr))
        out = np.zeros(self.n)
        np.add.at(out, rows, self.values * x[self.colind])
        return out


def forward(g, p, r, exact=False):
    """Returns z and the activations needed by backward."""
    f = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    dt = np.float64 if exact else np.float32
    rt = g.isq * r
    s = float(np.sqrt(np.mean(rt * rt)))
    s = s if s > 0 else 1.0
    u = (rt / s).astype(dt)
    W0 = p["W0"].astype(dt)
    pre0 = u[:, None] * W0[0][None, :] + W0[1][None, :]
    h = f(np.maximum(pre0, 0))
    acts = {"u": u, "s": s, "h0": h}
    for l in range(1, NLAYERS + 1):
        m = f(g.agg(h, dt))
        X = np.concatenate([h, m], axis=1)
        pre = X.astype(dt) @ f(p[f"W{l}"]).astype(dt) + p[f"b{l}"].astype(dt)
        h = f(np.maximum(pre, 0))
        acts[f"X{l}"] = X
        acts[f"h{l}"] = h
    y = h.astype(dt) @ p["w_out"].astype(dt) + dt(p["b_out"][0])
    acts["y"] = y
    z = g.isq * (rt + s * y.astype(np.float64))
    return z, acts


def backward(g, p, acts, dz, exact=False):
    """dL/dtheta for a given dL/dz (z = D^-1/2 (r~ + s y))."""
    f = (lambda a: np.asarray(a, np.float64)) if exact else bf16
    dt = np.float64 if exact else np.float32
    s = acts["s"]
    dy = (dz * g.isq * s).astype(dt)
    grads = {}
    h = acts[f"h{NLAYERS}"]
    grads["w_out"] = (h.astype(dt).T @ dy).astype(np.float64)
    grads["b_out"] = np.array([dy.sum(dtype=dt)], np.float64)
    dh = dy[:, None] * p["w_out"].astype(dt)[None, :]
    for l in range(NLAYERS, 0, -1):
        hl = acts[f"h{l}"]
        dpre = dh * (hl > 0)
        dpre_b = f(dpre)
        X = acts[f"X{l}"]
        grads[f"W{l}"] = (X.astype(dt).T @ dpre_b.astype(dt)).astype(np.float64)
        grads[f"b{l}"] = dpre.sum(axis=0, dtype=dt).astype(np.float64)
        dX = dpre_b.astype(dt) @ f(p[f"W{l}"]).astype(dt).T
        dh = dX[:, :C] + g.agg_t(dX[:, C:], dt)
    h0 = acts["h0"]
    dpre0 = dh * (h0 > 0)
    u = acts["u"]
    grads["W0"] = np.stack([(u[:, None] * dpre0).sum(axis=0, dtype=dt),
                            dpre0.sum(axis=0, dtype=dt)]).astype(np.float64)
    return grads


def loss_and_grad(g, p, bs, xs, exact=False):
    """Mean relative energy-norm error over the batch and its gradient."""
    tot = 0.0
    gsum = None
    B = len(bs)
    for b, x in zip(bs, xs):
        z, acts = forward(g, p, b, exact)
        e = z - x
        Ae = g.spmv(e)
        nx = float(x @ g.spmv(x))
        tot += float(e @ Ae) / nx / B
        dz = 2.0 * Ae / nx / B
        gr = backward(g, p, acts, dz, exact)
        gsum = gr if gsum is None else {k: gsum[k] + gr[k] for k in gsum}
    return tot, gsum


def adam_step(p, grads, state, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam update of the FP32 master weights (in place); state = {t, m, v}."""
    state["t"] = state.get("t", 0) + 1
    t = state["t"]
    for k in p:
        m = state.setdefault("m_" + k, np.zeros_like(p[k], dtype=np.float32))
        v = state.setdefault("v_" + k, np.zeros_like(p[k], dtype=np.float32))
        gk = grads[k].astype(np.float32)
        m[:] = b1 * m + (1 - b1) * gk
        v[:] = b2 * v + (1 - b2) * gk * gk
        mh = m / (1 - b1 ** t)
        vh = v / (1 - b2 ** t)
        p[k] = (p[k] - lr * mh / (np.sqrt(vh) + eps)).astype(np.float32)
    return p
