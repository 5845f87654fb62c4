"""The oracle's IC(0)/IC(p) (incomplete LDL^T, or_ic0_d) and preconditioned CG (or_cg_d),
which the reference announces but does not have, pinned on the CPU by what the
mathematics fixes:

* on a pattern with no fill (tridiagonal) the incomplete factorization is the complete
  one: L D L^T = A, and L sqrt(D) is numpy's Cholesky factor;
* on any symmetric pattern L D L^T reproduces A on the pattern (the defining property of
  IC(0), and of IC(p) on its fill pattern, with fill entries 0);
* for a symmetric matrix it equals the oracle's ILU(0) (pinned against the reference):
  the same L, and U = D L^T;
* CG on a matrix with k distinct eigenvalues converges in k steps, its solution is
  numpy's, and with the exact factorization as preconditioner it converges in one;
* planted mistakes (a dropped pivot factor d_m, an unsymmetric pattern) fail.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import dense, oracle_ilu0

o = O.orc()
P = O.P


def ic0(val, cind, rptr):
    f = np.ascontiguousarray(val, dtype=np.float64).copy()
    st = o.or_ic0_d(P(f), P(O.i32(cind)), P(O.i32(rptr)), len(rptr) - 1)
    return st, f


def ldl(f, cind, rptr):
    N = len(rptr) - 1
    F = dense(rptr, cind, f, N)
    L = np.tril(F, -1) + np.eye(N)
    d = np.diag(F).copy()
    return L, d, np.triu(F, 1)


def tridiag(n, seed=0):
    rng = np.random.default_rng(seed)
    off = rng.uniform(-1, 1, n - 1)
    A = np.diag(np.abs(np.concatenate([[0], off])) + np.abs(np.concatenate([off, [0]])) + rng.uniform(0.5, 1, n))
    A += np.diag(off, 1) + np.diag(off, -1)
    return A


def csr_of(A):
    rows, cols = np.nonzero(A)
    rptr = np.concatenate([[0], np.cumsum(np.bincount(rows, minlength=A.shape[0]))]).astype(np.int32)
    return A[rows, cols].astype(np.float64), cols.astype(np.int32), rptr


def test_no_fill_pattern_gives_the_complete_factorization():
    A = tridiag(40)
    val, cind, rptr = csr_of(A)
    st, f = ic0(val, cind, rptr)
    assert st == 0
    L, d, Up = ldl(f, cind, rptr)
    assert np.allclose(L @ np.diag(d) @ L.T, A, rtol=0, atol=1e-14 * np.abs(A).max())
    assert np.allclose(L * np.sqrt(d)[None, :], np.linalg.cholesky(A), rtol=1e-13, atol=1e-15)
    assert np.allclose(Up, np.triu(np.diag(d) @ L.T, 1), rtol=1e-15, atol=0)


@pytest.mark.parametrize("gen", ["poisson", "vardiff"])
def test_pattern_property_and_equality_with_ilu0(gen):
    rptr, cind, val = S.gen_poisson7(6) if gen == "poisson" else S.gen_vardiff7(6, 2, 3.0, 1.5)
    N = len(rptr) - 1
    A = dense(rptr, cind, val, N)
    assert np.array_equal(A, A.T)
    st, f = ic0(val, cind, rptr)
    assert st == 0
    L, d, _ = ldl(f, cind, rptr)
    R = L @ np.diag(d) @ L.T - A
    on = A != 0
    assert np.abs(R[on]).max() <= 1e-13 * np.abs(A).max()
    assert np.abs(R[~on]).max() > 1e-3  # the fill it drops is real
    g = oracle_ilu0(val, cind, rptr)  # the reference's IKJ ILU(0), pinned separately
    rows = np.repeat(np.arange(N), np.diff(rptr))
    assert np.allclose(f, g, rtol=1e-13, atol=1e-15 * np.abs(g).max())
    # the upper triangle is exactly D L^T (set from the lower one)
    F = dense(rptr, cind, f, N)
    assert np.array_equal(np.triu(F, 1), np.triu(np.diag(d) @ L.T, 1))
    assert rows.size == len(f)


def test_icp_on_the_level_fill_pattern():
    rptr, cind, val = S.gen_vardiff7(6, 4, 3.0, 1.5)
    N = len(rptr) - 1
    fv, fc, fr = S.alloc_level_zero(val, cind, rptr, N, 1)
    A = dense(rptr, cind, val, N)
    Fp = dense(fr, fc, np.ones(len(fc)), N) != 0
    assert np.array_equal(Fp, Fp.T) and Fp.sum() > (A != 0).sum()
    st, f = ic0(fv, fc, fr)
    assert st == 0
    L, d, _ = ldl(f, fc, fr)
    R = L @ np.diag(d) @ L.T - A
    assert np.abs(R[Fp]).max() <= 1e-13 * np.abs(A).max()
    # more fill, a better factor
    _, f0 = ic0(val, cind, rptr)
    L0, d0, _ = ldl(f0, cind, rptr)
    assert np.linalg.norm(R) < np.linalg.norm(L0 @ np.diag(d0) @ L0.T - A)


def test_errors_and_planted_mistake():
    A = tridiag(12, 3)
    val, cind, rptr = csr_of(A)
    # unsymmetric pattern: drop one upper entry
    keep = np.ones(len(cind), bool)
    keep[rptr[2] + 2] = False  # row 2's entry right of the diagonal
    v2, c2 = val[keep], cind[keep]
    r2 = np.concatenate([[0], np.cumsum(np.diff(rptr) - np.isin(np.arange(12), [2]))]).astype(np.int32)
    assert ic0(v2, c2, r2)[0] == -2
    # a factor with slightly wrong multipliers fails the closed form (the update's
    # l_im d_m l_km terms only arise with fill: test_icp_on_the_level_fill_pattern)
    st, f = ic0(val, cind, rptr)
    L, d, _ = ldl(f, cind, rptr)
    Lbad = L.copy()
    for i in range(1, 12):
        Lbad[i, i - 1] = A[i, i - 1] / d[i - 1] * (1 + 1e-3)  # perturbed multiplier
    assert not np.allclose(Lbad @ np.diag(d) @ Lbad.T, A, atol=1e-12)


def _cg(val, cind, rptr, b, prec=None, eps=1e-12, max_iter=500):
    N = len(rptr) - 1
    x = np.zeros(N)
    conv = C.c_int()
    hist = np.zeros(max_iter)
    spec = C.byref(prec) if prec is not None else None
    it = o.or_cg_d(P(val), P(O.i32(cind)), P(O.i32(rptr)), spec, P(b), P(x), float(np.linalg.norm(b)), max_iter, N,
                   eps, P(hist), max_iter, C.byref(conv))
    return it, conv.value, x, hist[:it]


def test_cg_closed_forms():
    rng = np.random.default_rng(1)
    # three distinct eigenvalues: three steps
    Q, _ = np.linalg.qr(rng.normal(size=(30, 30)))
    lam = np.repeat([1.0, 4.0, 9.0], 10)
    A = Q @ np.diag(lam) @ Q.T
    A = (A + A.T) / 2
    val, cind, rptr = csr_of(A)
    b = rng.normal(size=30)
    it, conv, x, _ = _cg(val, cind, rptr, b, eps=1e-10)
    assert conv == 1 and it == 3
    assert np.allclose(x, np.linalg.solve(A, b), rtol=1e-8)
    # an SPD stencil: numpy's solution, a monotone energy error
    rptr, cind, val = S.gen_vardiff7(5, 3, 3.0, 1.0)
    N = len(rptr) - 1
    A = dense(rptr, cind, val, N)
    b = rng.uniform(-1, 1, N)
    it, conv, x, hist = _cg(val, cind, rptr, b, eps=1e-11)
    assert conv == 1 and 0 < it < N
    assert np.linalg.norm(A @ x - b) <= 1.5e-11 * np.linalg.norm(b)
    # the exact factorization as preconditioner: one step (tridiagonal, no fill)
    A = tridiag(50, 5)
    val, cind, rptr = csr_of(A)
    st, f = ic0(val, cind, rptr)
    sp = S.split(f, cind, rptr, 50, "L-DU", True)
    spec = O.prec_spec(1, sp["L"], sp["U"])
    b = rng.uniform(-1, 1, 50)
    it, conv, x, _ = _cg(val, cind, rptr, b, prec=spec, eps=1e-12)
    assert conv == 1 and it == 1


def test_ic_preconditioned_cg_beats_plain_cg():
    rptr, cind, val = S.gen_vardiff7(8, 1, 4.0, 1.5)
    N = len(rptr) - 1
    b = np.random.default_rng(2).uniform(-1, 1, N)
    it0, c0, x0, _ = _cg(val, cind, rptr, b, eps=1e-10)
    st, f = ic0(val, cind, rptr)
    sp = S.split(f, cind, rptr, N, "L-DU", True)
    it1, c1, x1, _ = _cg(val, cind, rptr, b, prec=O.prec_spec(1, sp["L"], sp["U"]), eps=1e-10)
    assert c0 == 1 and c1 == 1 and it1 < 0.6 * it0
    A = dense(rptr, cind, val, N)
    assert np.linalg.norm(A @ x1 - b) <= 2e-10 * np.linalg.norm(b)
