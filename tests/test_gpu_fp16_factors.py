"""FP16 factors with a per-row scale, against an exact CPU emulation and the FP64 solve.

With the precision bit SENK_FACTOR_F16 a preconditioner keeps FP16 copies of its ILU
factor values in the solve layout: each row stores value * 2^-e, e the binary exponent
of the row's largest magnitude (the U diagonal slot excluded), the U inverted diagonal
stays FP32, and the arithmetic is FP32 in the order of the FP32 kernels.  Checks:
  * bit-exact against a numpy emulation of exactly that (FP16 rounding of the scaled
    values, then the sequential FP32 L and U solves in CSR order), in both launch modes;
  * the error against the FP64 application stays at FP16 resolution;
  * the row scale makes the stored values independent of the matrix's magnitude: for
    2^20 A (whose U values overflow FP16 unscaled) and 2^-30 A (whose values are below
    FP16's normal range) the application equals 2^-20 resp. 2^30 times that for A, bit
    for bit;
  * a nested mixed solve with FP16 factors in its inner FP32 solve still converges to
    the FP64 tolerance.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, relerr

pytestmark = pytest.mark.gpu

F16 = (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32) | (1 << S.SENK_FACTOR_F16)


def _handles(ctx, sy, scale=1.0):
    N = sy["N"]
    val = sy["val"] * scale
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(val), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, F16, 0)
    M = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(val),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts), C.byref(M))
    return A, M


def _download(M, N):
    ln, un = C.c_int(), C.c_int()
    S.call("senk_prec_factor_nnz", M, C.byref(ln), C.byref(un), None, None)
    lr, ur = np.zeros(N + 1, np.int32), np.zeros(N + 1, np.int32)
    lc, uc = np.zeros(ln.value, np.int32), np.zeros(un.value, np.int32)
    lv, uv = np.zeros(ln.value), np.zeros(un.value)
    S.call("senk_prec_download", M, O.P(lr), O.P(lc), O.P(lv), O.P(ur), O.P(uc), O.P(uv))
    return (lv, lc, lr), (uv, uc, ur)


def _quantise(v, first):
    """Row values from `first` on -> FP32 values 2^e * fp16(v * 2^-e)."""
    out = v.astype(np.float32).copy()
    w = v[first:]
    if w.size:
        m = np.max(np.abs(w))
        e = np.frexp(m)[1] if m > 0 else 0
        h = np.ldexp(w, -e).astype(np.float32).astype(np.float16)
        out[first:] = np.ldexp(h.astype(np.float32), e).astype(np.float32)
    return out


def _emulate(L, U, x):
    """Sequential FP32 solves with the FP16-quantised factors, CSR order."""
    lv, lc, lr = L
    uv, uc, ur = U
    N = len(x)
    y = np.zeros(N, np.float32)
    for i in range(N):
        q = _quantise(lv[lr[i]:lr[i + 1]], 0)
        t = np.float32(x[i])
        for k, j in enumerate(range(lr[i], lr[i + 1])):
            t = np.float32(t - np.float32(q[k] * y[lc[j]]))
        y[i] = t
    z = np.zeros(N, np.float32)
    for i in range(N - 1, -1, -1):
        s, e = ur[i], ur[i + 1]
        q = _quantise(uv[s:e], 1)
        t = np.float32(y[i])
        for j in range(e - 1, s, -1):
            t = np.float32(t - np.float32(q[j - s] * z[uc[j]]))
        z[i] = np.float32(t * np.float32(uv[s]))
    return z


def _apply16(ctx, M, x):
    dx = ctx.array(np.asarray(x, np.float32))
    dy = ctx.array(np.zeros(len(x), np.float32))
    S.call("senk_prec_apply_f16", M, dx.p, dy.p)
    return dy.download()


@pytest.mark.parametrize("gen,bsize", [("poisson", 4), ("vardiff", 8), ("convdiff", 16)])
def test_fp16_apply_matches_emulation(ctx, gen, bsize):
    if gen == "poisson":
        rptr, cind, val = S.gen_poisson7(14)
    elif gen == "vardiff":
        rptr, cind, val = S.gen_vardiff7(14, 3, 8.0, 1.5)
    else:
        rptr, cind, val = S.gen_convdiff27(12, 10.0)
    sy = abmc_system(rptr, cind, val, bsize)
    N = sy["N"]
    A, M = _handles(ctx, sy)
    try:
        L, U = _download(M, N)
        x = np.random.default_rng(1).uniform(-1, 1, N).astype(np.float32)
        want = _emulate(L, U, x)
        for mode in (0, 1, 2, 3):
            S.call("senk_prec_set_trsv_mode", M, mode)
            got = _apply16(ctx, M, x)
            assert np.array_equal(got, want), (mode, np.max(np.abs(got - want)))
        S.call("senk_prec_check", M)
        # against the FP64 factors: FP16 resolution (2^-11 relative per stored value)
        dx, dy = ctx.array(x.astype(np.float64)), ctx.zeros(N)
        S.call("senk_prec_apply_f64", M, dx.p, dy.p)
        err = relerr(want, dy.download())
        assert err <= 5e-3, err
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_fp16_row_scale_makes_magnitude_irrelevant(ctx):
    rptr, cind, val = S.gen_vardiff7(12, 5, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, 4)
    N = sy["N"]
    x = np.random.default_rng(2).uniform(-1, 1, N).astype(np.float32)
    base = None
    for k in (0, 20, -30):
        A, M = _handles(ctx, sy, 2.0 ** k)
        try:
            if k == 20:
                # unscaled FP16 storage of these U values would overflow
                _, (uv, _, ur) = _download(M, N)
                offdiag = np.delete(uv, ur[:-1])
                assert np.max(np.abs(offdiag)) > 65504
            z = _apply16(ctx, M, x)
            if base is None:
                base = z
            else:
                assert np.array_equal(z, np.ldexp(base, -k).astype(np.float32)), k
        finally:
            S.lib().senk_prec_destroy(M)
            S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("solver", [S.SENK_BICGSTAB, S.SENK_FGCR])
def test_nested_mixed_with_fp16_factors(ctx, solver):
    rptr, cind, val = S.gen_vardiff7(24, 3, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, 8)
    N = sy["N"]
    xs = np.random.default_rng(4).uniform(-1, 1, N)
    rows = np.repeat(np.arange(N), np.diff(sy["rptr"]))
    b = np.bincount(rows, weights=sy["val"] * xs[sy["cind"]], minlength=N)
    A, M = _handles(ctx, sy)
    try:
        db = ctx.array(b)
        its = {}
        for fp in (S.SENK_FACTOR_F32, S.SENK_FACTOR_F16):
            dx = ctx.zeros(N)
            o = S.solve_opts(solver=solver, m=30, eps=1e-10, mixed=2, inner_eps=1e-3,
                             inner_max=40 if solver == S.SENK_BICGSTAB else 30, outer_m=20, max_iter=6000,
                             factor_prec=fp)
            res = S.senk_result()
            S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0)
            assert res.converged == 1 and res.status == 0, fp
            assert res.true_relres <= 1e-10 * 1.5, (fp, res.true_relres)
            assert relerr(dx.download(), xs) <= 1e-5, fp
            its[fp] = res.iters
        # FP16 factors cost at most a quarter more iterations here
        assert its[S.SENK_FACTOR_F16] <= 1.25 * its[S.SENK_FACTOR_F32] + 2, its
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_fp16_needs_the_layout(ctx):
    rptr, cind, val = S.gen_poisson7(8)
    sy = abmc_system(rptr, cind, val, 4)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    try:
        opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, F16, 1)  # csr_only
        M = C.c_void_p()
        st = S.lib().senk_prec_create_ilu(ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
                                          sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], 4, C.byref(opts),
                                          C.byref(M))
        assert st == 9  # SENK_ERR_UNSUPPORTED
    finally:
        S.lib().senk_mat_destroy(A)
