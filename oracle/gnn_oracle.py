"""CPU restatement (numpy, FP64) of the learned ILU correction: forward pass, loss and
backward pass.  Test infrastructure only -- the product's network runs in
senk_amd/csrc/gnn.hip; this file is the checker its kernels are compared against.

The network has no counterpart in the reference (it is a new module beside ILU); this
file is its definition.  Per application, for one residual r with ILU output
z0 = M^{-1} r on a matrix A with diagonal d:

    s      = |d|^(-1/2)                   Ahat = S A S  (S = diag(s))
    rh     = s * r,  zh = z0 / s          sigma = rms(rh)   (no correction if 0)
    x      = [rh/sigma, zh/sigma]         (N x 2)
    h1     = relu(x  W1s + (Ahat x)  W1n + b1)      (N x 32)
    h2     = relu(h1 W2s + (Ahat h1) W2n + b2)
    h3     = relu(h2 W3s + (Ahat h2) W3n + b3)
    chat   = h3 wo + bo                    c = sigma * s * chat
    z      = z0 + c

Loss over a batch of residuals on one matrix:  L = mean_b ||A z_b - r_b||^2 / ||r_b||^2.

Parameters are one flat FP32 vector in the order of PARAM_SHAPES.  The device keeps
BF16 copies of the weights; with emulate_bf16=True the oracle rounds the same tensors
the device rounds (weights, the stored activations h1..h3, and the aggregated inputs of
the matrix-core layers Ahat h1, Ahat h2) so that the two agree to FP32 accumulation
order.

Pinning: the backward pass is checked against central finite differences of this
forward pass (tests/test_gnn_oracle.py); the forward pass against a direct dense
evaluation of the formulas above.  Parity of the oracle with itself is therefore
fixed by calculus, not by a second transcription.
"""
from __future__ import annotations

import numpy as np

WIDTH = 32
F_IN = 2
PARAM_SHAPES = [
    ("W1s", (F_IN, WIDTH)), ("W1n", (F_IN, WIDTH)), ("b1", (WIDTH,)),
    ("W2s", (WIDTH, WIDTH)), ("W2n", (WIDTH, WIDTH)), ("b2", (WIDTH,)),
    ("W3s", (WIDTH, WIDTH)), ("W3n", (WIDTH, WIDTH)), ("b3", (WIDTH,)),
    ("wo", (WIDTH,)), ("bo", (1,)),
]
N_PARAMS = sum(int(np.prod(s)) for _, s in PARAM_SHAPES)
BF16_ROUNDED = {"W1s", "W1n", "W2s", "W2n", "W3s", "W3n", "wo"}


def bf16(a):
    """Round FP32 values to the nearest BF16 (ties to even), returned as float64."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def unpack(theta):
    theta = np.asarray(theta, dtype=np.float64)
    assert theta.size == N_PARAMS
    out, o = {}, 0
    for name, shp in PARAM_SHAPES:
        n = int(np.prod(shp))
        out[name] = theta[o:o + n].reshape(shp)
        o += n
    return out


def pack(p):
    return np.concatenate([np.asarray(p[name], dtype=np.float64).ravel() for name, _ in PARAM_SHAPES])


def init_params(seed=0):
    """He-style initialisation; the output layer starts at zero (no correction)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shp in PARAM_SHAPES:
        if name.startswith("W"):
            fan_in = shp[0] * 2
            p[name] = rng.normal(0, np.sqrt(2.0 / fan_in), shp)
        else:
            p[name] = np.zeros(shp)
    return pack(p).astype(np.float32).astype(np.float64)


def _csr_dense(rptr, cind, val, N):
    A = np.zeros((N, N))
    for i in range(N):
        for j in range(rptr[i], rptr[i + 1]):
            A[i, cind[j]] += val[j]
    return A


class Graph:
    def __init__(self, rptr, cind, val):
        self.N = len(rptr) - 1
        self.A = _csr_dense(rptr, cind, val, self.N)
        d = np.diag(self.A).copy()
        self.s = 1.0 / np.sqrt(np.abs(d))
        self.Ah = self.s[:, None] * self.A * self.s[None, :]


def forward(theta, g: Graph, r, z0, emulate_bf16=False, keep=False):
    p = unpack(theta)
    if emulate_bf16:
        p = {k: (bf16(v) if k in BF16_ROUNDED else v) for k, v in p.items()}
    rnd = bf16 if emulate_bf16 else (lambda a: a)
    rh = g.s * r
    zh = z0 / g.s
    sigma = np.sqrt(np.mean(rh * rh))
    if sigma == 0:
        return (z0.copy(), {}) if keep else z0.copy()
    x = np.stack([rh / sigma, zh / sigma], axis=1)
    m1 = g.Ah @ x
    h1 = rnd(np.maximum(x @ p["W1s"] + m1 @ p["W1n"] + p["b1"], 0))
    m2 = rnd(g.Ah @ h1)
    h2 = rnd(np.maximum(h1 @ p["W2s"] + m2 @ p["W2n"] + p["b2"], 0))
    m3 = rnd(g.Ah @ h2)
    h3 = rnd(np.maximum(h2 @ p["W3s"] + m3 @ p["W3n"] + p["b3"], 0))
    chat = h3 @ p["wo"] + p["bo"][0]
    c = sigma * g.s * chat
    z = z0 + c
    if keep:
        return z, dict(p=p, x=x, m1=m1, h1=h1, m2=m2, h2=h2, m3=m3, h3=h3, sigma=sigma)
    return z


def loss(theta, g: Graph, R, Z0, emulate_bf16=False):
    tot = 0.0
    for r, z0 in zip(R, Z0):
        z = forward(theta, g, r, z0, emulate_bf16)
        e = g.A @ z - r
        tot += (e @ e) / (r @ r)
    return tot / len(R)


def _agg_T(g: Graph, v):
    """Backward of the aggregation m = Ahat h: dh += Ahat^T dm."""
    return g.Ah.T @ v


def loss_grad(theta, g: Graph, R, Z0, emulate_bf16=False):
    """Returns (L, dL/dtheta) by reverse-mode differentiation of forward()."""
    grads = {name: np.zeros(shp) for name, shp in PARAM_SHAPES}
    tot = 0.0
    B = len(R)
    for r, z0 in zip(R, Z0):
        z, t = forward(theta, g, r, z0, emulate_bf16, keep=True)
        e = g.A @ z - r
        rr = r @ r
        tot += (e @ e) / rr
        if not t:
            continue
        p = t["p"]
        dz = (2.0 / B) * (g.A.T @ e) / rr
        dchat = dz * t["sigma"] * g.s
        grads["wo"] += t["h3"].T @ dchat
        grads["bo"] += dchat.sum(keepdims=True)
        dh = np.outer(dchat, p["wo"])
        for l, (hin, mag, hout) in ((3, (t["h2"], t["m3"], t["h3"])), (2, (t["h1"], t["m2"], t["h2"])),
                                    (1, (t["x"], t["m1"], t["h1"]))):
            gl = dh * (hout > 0)
            grads[f"W{l}s"] += hin.T @ gl
            grads[f"W{l}n"] += mag.T @ gl
            grads[f"b{l}"] += gl.sum(axis=0)
            if l > 1:
                dh = gl @ p[f"W{l}s"].T + _agg_T(g, gl @ p[f"W{l}n"].T)
    return tot / B, pack(grads)
