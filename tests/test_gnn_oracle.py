"""Pins of the learned correction's CPU oracle (oracle/gnn_oracle.py).

The network is this build's own definition, so the oracle is pinned by mathematics:
  * the reverse-mode gradient equals central finite differences of the forward pass;
  * the forward pass equals a direct evaluation of the stated formulas (written with
    explicit loops over nodes and edges, not matrix products);
  * the BF16 rounding equals torch's float32 -> bfloat16 conversion (round to nearest
    even);
  * the correction is homogeneous (scaling r and z0 scales it) and vanishes at the
    initial parameters (output layer zero), so the untrained network is plain ILU.
"""
import numpy as np
import pytest

from oracle import gnn_oracle as G

import senk_amd as S


def _graph(n=4, seed=3):
    rptr, cind, val = S.gen_vardiff7(n, seed, 2.0, 1.0)
    return G.Graph(rptr, cind, val), (rptr, cind, val)


def _params(seed=0, scale=0.3):
    rng = np.random.default_rng(seed)
    th = G.init_params(seed)
    p = G.unpack(th)
    p = {k: v.copy() for k, v in p.items()}
    p["wo"] = rng.normal(0, scale, p["wo"].shape)
    p["bo"] = np.array([0.05])
    for b in ("b1", "b2", "b3"):
        p[b] = rng.normal(0, 0.1, p[b].shape)
    return G.pack(p)


def test_gradient_matches_finite_differences():
    g, _ = _graph(4)
    rng = np.random.default_rng(1)
    R = [rng.normal(size=g.N) for _ in range(2)]
    Z0 = [np.linalg.solve(np.tril(g.A) + 0.5 * np.eye(g.N), r) for r in R]
    th = _params(2)
    L, grad = G.loss_grad(th, g, R, Z0)
    assert np.isclose(L, G.loss(th, g, R, Z0), rtol=1e-12)
    idx = rng.choice(G.N_PARAMS, 60, replace=False)
    idx = np.concatenate([idx, [G.N_PARAMS - 1, 0, 64, 128]])  # bo, W1s, W1n, b1
    h = 1e-6
    for i in idx:
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss(tp, g, R, Z0) - G.loss(tm, g, R, Z0)) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 + 1e-5 * abs(fd), (i, fd, grad[i])


def test_forward_matches_loop_evaluation():
    g, (rptr, cind, val) = _graph(3)
    N = g.N
    rng = np.random.default_rng(5)
    r, z0 = rng.normal(size=N), rng.normal(size=N)
    th = _params(4)
    p = G.unpack(th)
    d = np.array([val[[j for j in range(rptr[i], rptr[i + 1]) if cind[j] == i][0]] for i in range(N)])
    s = 1 / np.sqrt(np.abs(d))
    sig = np.sqrt(np.mean((s * r) ** 2))
    x = [np.array([s[i] * r[i], z0[i] / s[i]]) / sig for i in range(N)]

    def agg(h):
        out = []
        for i in range(N):
            acc = np.zeros_like(h[0])
            for j in range(rptr[i], rptr[i + 1]):
                acc = acc + s[i] * val[j] * s[cind[j]] * h[cind[j]]
            out.append(acc)
        return out

    def layer(h, Ws, Wn, b):
        m = agg(h)
        return [np.maximum(h[i] @ Ws + m[i] @ Wn + b, 0) for i in range(N)]

    h1 = layer(x, p["W1s"], p["W1n"], p["b1"])
    h2 = layer(h1, p["W2s"], p["W2n"], p["b2"])
    h3 = layer(h2, p["W3s"], p["W3n"], p["b3"])
    c = np.array([sig * s[i] * (h3[i] @ p["wo"] + p["bo"][0]) for i in range(N)])
    assert np.allclose(G.forward(th, g, r, z0), z0 + c, rtol=1e-12, atol=1e-12)


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(0)
    a = np.concatenate([rng.normal(size=5000).astype(np.float32),
                        np.array([1.0, -1.0, 0.0, 1 + 2 ** -8, 1 + 3 * 2 ** -8, 3.0039062], np.float32)])
    want = torch.from_numpy(a).to(torch.bfloat16).to(torch.float32).numpy().astype(np.float64)
    assert np.array_equal(G.bf16(a), want)


def test_homogeneous_and_identity_at_init():
    g, _ = _graph(4)
    rng = np.random.default_rng(2)
    r, z0 = rng.normal(size=g.N), rng.normal(size=g.N)
    th = _params(1)
    z = G.forward(th, g, r, z0)
    for a in (3.0, 1e-3):
        assert np.allclose(G.forward(th, g, a * r, a * z0), a * z, rtol=1e-10, atol=1e-14)
    assert np.array_equal(G.forward(G.init_params(0), g, r, z0), z0)


def test_planted_mistake_is_caught(monkeypatch):
    # plant a classic mistake in the oracle's backward pass -- aggregating with Ahat
    # instead of Ahat^T -- and check that the finite-difference pin rejects it on a
    # nonsymmetric matrix (while the correct backward passes it)
    rptr, cind, val = S.gen_unstructured(3000, 2)
    N = 60
    # leading 60x60 block of the nonsymmetric generator output (columns < 60 kept)
    rows = [[(cind[j], val[j]) for j in range(rptr[i], rptr[i + 1]) if cind[j] < N] for i in range(N)]
    rp = np.cumsum([0] + [len(r) for r in rows]).astype(np.int32)
    ci = np.array([c for r in rows for c, _ in r], np.int32)
    va = np.array([v for r in rows for _, v in r])
    g = G.Graph(rp, ci, va)
    assert not np.allclose(g.Ah, g.Ah.T)
    rng = np.random.default_rng(3)
    R, Z0 = [rng.normal(size=N)], [rng.normal(size=N)]
    th = _params(3, scale=1.0)
    _, good = G.loss_grad(th, g, R, Z0)
    monkeypatch.setattr(G, "_agg_T", lambda gg, v: gg.Ah @ v)
    _, bad = G.loss_grad(th, g, R, Z0)
    monkeypatch.undo()
    h = 1e-6
    worst_good = worst_bad = 0.0
    for i in range(G.N_PARAMS):
        if abs(good[i] - bad[i]) < 1e-9:
            continue
        tp, tm = th.copy(), th.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss(tp, g, R, Z0) - G.loss(tm, g, R, Z0)) / (2 * h)
        worst_good = max(worst_good, abs(fd - good[i]) / (1e-6 + abs(fd)))
        worst_bad = max(worst_bad, abs(fd - bad[i]) / (1e-6 + abs(fd)))
        if worst_bad > 1e-2:
            break
    assert worst_good < 1e-4
    assert worst_bad > 1e-2
