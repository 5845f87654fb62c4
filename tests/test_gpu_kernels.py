"""GPU parity of the sparse kernels, BLAS-1 and ILU(0) against the oracle.

Tolerances: the order-preserving kernels (SpTRSV, ILU(0), SELL SpMV, element-wise
BLAS-1) must match the oracle bit for bit (both run without contraction in the
reference's operation order); kernels that split a row over lanes (CSR SpMV, BCSR
SpMV, the any-size SpTRSV fallback) must match to 1e-12 relative in the 2-norm.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, dense, ilu_factors, oracle_ilu0, relerr

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _mats():
    yield "poisson7_12", S.gen_poisson7(12)
    yield "convdiff27_10", S.gen_convdiff27(10, 10.0)
    yield "unstructured_3000", S.gen_unstructured(3000, 3)
    yield "vardiff7_10", S.gen_vardiff7(10, 2, 4.0, 1.5)


@pytest.mark.parametrize("name,mat", list(_mats()))
def test_spmv_csr_f64(ctx, name, mat):
    rptr, cind, val = mat
    N = len(rptr) - 1
    x = np.random.default_rng(1).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(x), O.P(yo), N)
    dv, dc, dr, dx, dy = ctx.array(val), ctx.array(cind), ctx.array(rptr), ctx.array(x), ctx.zeros(N)
    S.call("senk_spmv_csr_f64", ctx.h, dv.p, dc.p, dr.p, dx.p, dy.p, N)
    y = dy.download()
    assert relerr(y, yo) <= TOL, name
    # dense brute force
    if N <= 2000:
        assert relerr(y, dense(rptr, cind, val, N) @ x) <= 1e-12


def test_spmv_csr_f32_and_diag(ctx):
    rptr, cind, val = S.gen_convdiff27(9)
    N = len(rptr) - 1
    x = np.random.default_rng(2).uniform(-1, 1, N)
    v32, x32 = val.astype(np.float32), x.astype(np.float32)
    yo = np.zeros(N, np.float32)
    O.orc().or_spmv_csr_f(O.P(v32), O.P(cind), O.P(rptr), O.P(x32), O.P(yo), N)
    dv, dc, dr, dx, dy = ctx.array(v32), ctx.array(cind), ctx.array(rptr), ctx.array(x32), ctx.zeros(N, np.float32)
    S.call("senk_spmv_csr_f32", ctx.h, dv.p, dc.p, dr.p, dx.p, dy.p, N)
    assert relerr(dy.download(), yo) <= 1e-6
    # diag form: off-diagonal CSR + separate diagonal
    sp = S.split(val, cind, rptr, N, "LU-D", False)
    lv, lc, lr = sp["L"]
    d = sp["diag"]
    yo = np.zeros(N)
    O.orc().or_spmv_csr_diag_d(O.P(lv), O.P(lc), O.P(lr), O.P(d), O.P(x), O.P(yo), N)
    a = [ctx.array(t) for t in (lv, lc, lr, d, x)]
    dy = ctx.zeros(N)
    S.call("senk_spmv_csr_diag_f64", ctx.h, *(t.p for t in a), dy.p, N)
    assert np.array_equal(dy.download(), yo)


@pytest.mark.parametrize("N_extra", [0, 37])
def test_spmv_sell_reference_layout(ctx, N_extra):
    rptr, cind, val = S.gen_unstructured(64 * 40 + N_extra, 5)
    N = len(rptr) - 1
    sv, sc, sl = S.csr2sell(val, cind, rptr, 64, N)
    x = np.random.default_rng(3).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.orc().or_spmv_sell_d(O.P(sv), O.P(sc), O.P(sl), 64, O.P(x), O.P(yo), N)
    # the SELL sum order is the CSR row order: equal to the CSR oracle too
    yc = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(x), O.P(yc), N)
    assert np.array_equal(yo, yc)
    a = [ctx.array(t) for t in (sv, sc, sl)]
    dx, dy = ctx.array(x), ctx.zeros(N)
    S.call("senk_spmv_sell_f64", ctx.h, a[0].p, a[1].p, a[2].p, 64, dx.p, dy.p, N)
    assert np.array_equal(dy.download(), yo)


@pytest.mark.parametrize("sigma", [1, 64, 1024])
def test_spmv_sellcs(ctx, sigma):
    rptr, cind, val = S.gen_unstructured(5000, 7)
    N = len(rptr) - 1
    h = S.csr2sellcs(val, cind, rptr, N, 64, sigma)
    d = S.senk_sellcs()
    S.call("senk_sellcs_upload", ctx.h, C.byref(h), 1, C.byref(d))
    x = np.random.default_rng(4).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(x), O.P(yo), N)
    dx, dy = ctx.array(x), ctx.zeros(N)
    S.call("senk_spmv_sellcs_f64", ctx.h, C.byref(d), dx.p, dy.p)
    assert np.array_equal(dy.download(), yo)
    dx32, dy32 = ctx.array(x.astype(np.float32)), ctx.zeros(N, np.float32)
    S.call("senk_spmv_sellcs_f32", ctx.h, C.byref(d), dx32.p, dy32.p)
    assert relerr(dy32.download(), yo) <= 1e-5
    S.call("senk_sellcs_free_device", ctx.h, C.byref(d))
    S.lib().senk_sellcs_free_host(C.byref(h))


@pytest.mark.parametrize("bnl,bnw", [(3, 3), (8, 1), (2, 2), (4, 2)])
def test_spmv_bcsr(ctx, bnl, bnw):
    rptr, cind, val = S.gen_elasticity3(6, 3) if (bnl, bnw) == (3, 3) else S.gen_poisson7(8)
    N = len(rptr) - 1
    bv, bc, br = S.csr2bcsr(val, cind, rptr, N, bnl, bnw)
    x = np.random.default_rng(5).uniform(-1, 1, N)
    yo = np.zeros(N)
    O.orc().or_spmv_bcsr_d(O.P(bv), O.P(bc), O.P(br), O.P(x), O.P(yo), N, bnl, bnw)
    a = [ctx.array(t) for t in (bv, bc, br)]
    dx, dy = ctx.array(x), ctx.zeros(N)
    S.call("senk_spmv_bcsr_f64", ctx.h, a[0].p, a[1].p, a[2].p, dx.p, dy.p, N, bnl, bnw)
    assert relerr(dy.download(), yo) <= TOL


def _abmc_factors(kind, bsize):
    if kind == "poisson":
        rptr, cind, val = S.gen_poisson7(12)
    else:
        rptr, cind, val = S.gen_convdiff27(8, 10.0)
    sy = abmc_system(rptr, cind, val, bsize)
    L, U = ilu_factors(sy)
    return sy, L, U


@pytest.mark.parametrize("kind,bsize", [("poisson", 4), ("poisson", 16), ("convdiff", 8), ("convdiff", 32),
                                        ("poisson", 64), ("convdiff", 128), ("poisson", 48)])
def test_sptrsv_abmc(ctx, kind, bsize):
    sy, L, U = _abmc_factors(kind, bsize)
    N, cptr, nc = sy["N"], sy["cptr"], sy["nc"]
    x = np.random.default_rng(6).uniform(-1, 1, N)
    yl, yu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_l_d(O.P(L[0]), O.P(L[1]), O.P(L[2]), O.P(x), O.P(yl), N)
    O.orc().or_sptrsv_u_d(O.P(U[0]), O.P(U[1]), O.P(U[2]), O.P(yl), O.P(yu), N)
    # colored == sequential on the reordered factors (the oracle's own pin)
    ya = np.zeros(N)
    O.orc().or_sptrsv_l_abmc_d(O.P(L[0]), O.P(L[1]), O.P(L[2]), O.P(x), O.P(ya), N, O.P(cptr), nc, bsize)
    assert np.array_equal(ya, yl)
    dL = [ctx.array(t) for t in L]
    dU = [ctx.array(t) for t in U]
    dx, dt = ctx.array(x), ctx.zeros(N)
    cp = cptr.ctypes.data_as(C.c_void_p)
    S.call("senk_sptrsv_csr_l_abmc_f64", ctx.h, dL[0].p, dL[1].p, dL[2].p, dx.p, dt.p, N, cp, nc, bsize)
    t = dt.download()
    S.call("senk_sptrsv_csr_u_abmc_f64", ctx.h, dU[0].p, dU[1].p, dU[2].p, dt.p, dt.p, N, cp, nc, bsize)
    u = dt.download()
    if 64 % bsize == 0:
        assert np.array_equal(t, yl)
        assert np.array_equal(u, yu)
    else:
        assert relerr(t, yl) <= TOL and relerr(u, yu) <= TOL
    # residual pin independent of the oracle's loops: L t = x, U u = t
    Lm = np.eye(N) + dense(L[2], L[1], L[0], N)
    assert np.linalg.norm(Lm @ yl - x) <= 1e-10 * np.linalg.norm(x)


def test_sptrsv_abmc_f32(ctx):
    sy, L, U = _abmc_factors("convdiff", 16)
    N, cptr, nc = sy["N"], sy["cptr"], sy["nc"]
    x = np.random.default_rng(7).uniform(-1, 1, N).astype(np.float32)
    L32 = (L[0].astype(np.float32), L[1], L[2])
    U32 = (U[0].astype(np.float32), U[1], U[2])
    yl, yu = np.zeros(N, np.float32), np.zeros(N, np.float32)
    O.orc().or_sptrsv_l_f(O.P(L32[0]), O.P(L32[1]), O.P(L32[2]), O.P(x), O.P(yl), N)
    O.orc().or_sptrsv_u_f(O.P(U32[0]), O.P(U32[1]), O.P(U32[2]), O.P(yl), O.P(yu), N)
    dL = [ctx.array(t) for t in L32]
    dU = [ctx.array(t) for t in U32]
    dx, dt = ctx.array(x), ctx.zeros(N, np.float32)
    cp = cptr.ctypes.data_as(C.c_void_p)
    S.call("senk_sptrsv_csr_l_abmc_f32", ctx.h, dL[0].p, dL[1].p, dL[2].p, dx.p, dt.p, N, cp, nc, 16)
    S.call("senk_sptrsv_csr_u_abmc_f32", ctx.h, dU[0].p, dU[1].p, dU[2].p, dt.p, dt.p, N, cp, nc, 16)
    assert np.array_equal(dt.download(), yu)


def test_sptrsv_amc(ctx):
    rptr, cind, val = S.gen_convdiff27(8)
    N = len(rptr) - 1
    nc, sc, LP, RP = S.amc_permutation(cind, rptr, N, False)
    v2, c2, r2 = S.reordering(val, cind, rptr, LP, RP, N)
    L, U = ilu_factors(dict(val=v2, cind=c2, rptr=r2))
    x = np.random.default_rng(8).uniform(-1, 1, N)
    yl, yu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_l_amc_d(O.P(L[0]), O.P(L[1]), O.P(L[2]), O.P(x), O.P(yl), N, O.P(sc), nc)
    O.orc().or_sptrsv_u_amc_d(O.P(U[0]), O.P(U[1]), O.P(U[2]), O.P(yl), O.P(yu), N, O.P(sc), nc)
    dL = [ctx.array(t) for t in L]
    dU = [ctx.array(t) for t in U]
    dx, dt = ctx.array(x), ctx.zeros(N)
    cp = sc.ctypes.data_as(C.c_void_p)
    S.call("senk_sptrsv_csr_l_amc_f64", ctx.h, dL[0].p, dL[1].p, dL[2].p, dx.p, dt.p, N, cp, nc)
    S.call("senk_sptrsv_csr_u_amc_f64", ctx.h, dU[0].p, dU[1].p, dU[2].p, dt.p, dt.p, N, cp, nc)
    assert np.array_equal(dt.download(), yu)


@pytest.mark.parametrize("bj", [0, 4])
def test_sptrsv_sequential_and_block_jacobi(ctx, bj):
    rptr, cind, val = S.gen_poisson7(10)
    N = len(rptr) - 1
    if bj:
        val, cind, rptr = S.remove_off_diagonal(val, cind, rptr, N, bj)
    L, U = ilu_factors(dict(val=val, cind=cind, rptr=rptr))
    x = np.random.default_rng(9).uniform(-1, 1, N)
    yl, yu = np.zeros(N), np.zeros(N)
    if bj:
        O.orc().or_sptrsv_l_bj_d(O.P(L[0]), O.P(L[1]), O.P(L[2]), O.P(x), O.P(yl), N, bj)
        O.orc().or_sptrsv_u_bj_d(O.P(U[0]), O.P(U[1]), O.P(U[2]), O.P(yl), O.P(yu), N, bj)
    else:
        O.orc().or_sptrsv_l_d(O.P(L[0]), O.P(L[1]), O.P(L[2]), O.P(x), O.P(yl), N)
        O.orc().or_sptrsv_u_d(O.P(U[0]), O.P(U[1]), O.P(U[2]), O.P(yl), O.P(yu), N)
    lp, up = C.c_void_p(), C.c_void_p()
    S.call("senk_trsv_plan_create", ctx.h, O.P(L[2]), O.P(L[1]), N, 0, C.byref(lp))
    S.call("senk_trsv_plan_create", ctx.h, O.P(U[2]), O.P(U[1]), N, 1, C.byref(up))
    dL = [ctx.array(t) for t in L]
    dU = [ctx.array(t) for t in U]
    dx, dt = ctx.array(x), ctx.zeros(N)
    if bj:
        S.call("senk_sptrsv_csr_l_bj_f64", ctx.h, lp, dL[0].p, dL[1].p, dL[2].p, dx.p, dt.p, N, bj)
        S.call("senk_sptrsv_csr_u_bj_f64", ctx.h, up, dU[0].p, dU[1].p, dU[2].p, dt.p, dt.p, N, bj)
    else:
        S.call("senk_sptrsv_csr_l_f64", ctx.h, lp, dL[0].p, dL[1].p, dL[2].p, dx.p, dt.p, N)
        S.call("senk_sptrsv_csr_u_f64", ctx.h, up, dU[0].p, dU[1].p, dU[2].p, dt.p, dt.p, N)
    assert np.array_equal(dt.download(), yu)
    S.lib().senk_trsv_plan_destroy(lp)
    S.lib().senk_trsv_plan_destroy(up)


@pytest.mark.parametrize("bnl,bnw,bsize", [(3, 3, 12), (8, 1, 16), (2, 2, 8)])
def test_sptrsv_bcsr(ctx, bnl, bnw, bsize):
    if bnl == 3:
        rptr, cind, val = S.gen_elasticity3(4, 2)
    else:
        rptr, cind, val = S.gen_poisson7(8)
    N = len(rptr) - 1
    sy = abmc_system(rptr, cind, val, bsize)
    fv, fc, fr = S.alloc_block_zero(sy["val"], sy["cind"], sy["rptr"], N, bnl, bnw)
    f = oracle_ilu0(fv, fc, fr)
    sp = S.split(f, fc, fr, N, "L-DU", True)
    bl = S.csr2bcsr(*sp["L"], N, bnl, bnw)
    bu = S.csr2bcsr(*sp["U"], N, bnl, bnw)
    x = np.random.default_rng(10).uniform(-1, 1, N)
    cptr, nc = sy["cptr"], sy["nc"]
    yl, yu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_bcsr_l_abmc_d(O.P(bl[0]), O.P(bl[1]), O.P(bl[2]), O.P(x), O.P(yl), N, O.P(cptr), nc, bsize,
                                   bnl, bnw)
    O.orc().or_sptrsv_bcsr_u_abmc_d(O.P(bu[0]), O.P(bu[1]), O.P(bu[2]), O.P(yl), O.P(yu), N, O.P(cptr), nc, bsize,
                                   bnl, bnw)
    # ILUB BCSR solve == CSR solve on the block-closed pattern
    sl, su = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_l_d(O.P(sp["L"][0]), O.P(sp["L"][1]), O.P(sp["L"][2]), O.P(x), O.P(sl), N)
    O.orc().or_sptrsv_u_d(O.P(sp["U"][0]), O.P(sp["U"][1]), O.P(sp["U"][2]), O.P(sl), O.P(su), N)
    assert relerr(yu, su) <= 1e-13
    dl = [ctx.array(t) for t in bl]
    du = [ctx.array(t) for t in bu]
    dx, dt = ctx.array(x), ctx.zeros(N)
    cp = cptr.ctypes.data_as(C.c_void_p)
    S.call("senk_sptrsv_bcsr_l_abmc_f64", ctx.h, dl[0].p, dl[1].p, dl[2].p, dx.p, dt.p, N, cp, nc, bsize, bnl, bnw)
    assert np.array_equal(dt.download(), yl)
    S.call("senk_sptrsv_bcsr_u_abmc_f64", ctx.h, du[0].p, du[1].p, du[2].p, dt.p, dt.p, N, cp, nc, bsize, bnl, bnw)
    assert np.array_equal(dt.download(), yu)
    # sequential BCSR forms
    ql, qu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_bcsr_l_d(O.P(bl[0]), O.P(bl[1]), O.P(bl[2]), O.P(x), O.P(ql), N, bnl, bnw)
    O.orc().or_sptrsv_bcsr_u_d(O.P(bu[0]), O.P(bu[1]), O.P(bu[2]), O.P(ql), O.P(qu), N, bnl, bnw)
    dt2 = ctx.zeros(N)
    S.call("senk_sptrsv_bcsr_l_f64", ctx.h, dl[0].p, dl[1].p, dl[2].p, dx.p, dt2.p, N, bnl, bnw)
    S.call("senk_sptrsv_bcsr_u_f64", ctx.h, du[0].p, du[1].p, du[2].p, dt2.p, dt2.p, N, bnl, bnw)
    assert np.array_equal(dt2.download(), qu)


@pytest.mark.parametrize("kind,bsize", [("poisson", 4), ("convdiff", 16), ("convdiff", 128)])
def test_ilu0_device_colored(ctx, kind, bsize):
    sy, L, U = _abmc_factors(kind, bsize)
    N = sy["N"]
    fo = oracle_ilu0(sy["val"], sy["cind"], sy["rptr"])
    dv, dc, dr = ctx.array(sy["val"]), ctx.array(sy["cind"]), ctx.array(sy["rptr"])
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_ilu0_f64", ctx.h, dv.p, dc.p, dr.p, N, cp, sy["nc"], bsize, O.P(sy["rptr"]), O.P(sy["cind"]))
    assert np.array_equal(dv.download(), fo)
    # defining property (L U)_ij = A_ij on the pattern
    if N <= 1800:
        Lm = np.eye(N) + np.tril(dense(sy["rptr"], sy["cind"], fo, N), -1)
        Um = np.triu(dense(sy["rptr"], sy["cind"], fo, N))
        A = dense(sy["rptr"], sy["cind"], sy["val"], N)
        mask = A != 0
        assert np.max(np.abs((Lm @ Um - A)[mask])) <= 1e-12 * np.max(np.abs(A))


def test_ilu0_device_levels_and_f32(ctx):
    rptr, cind, val = S.gen_unstructured(2000, 11)
    N = len(rptr) - 1
    fo = oracle_ilu0(val, cind, rptr)
    dv, dc, dr = ctx.array(val), ctx.array(cind), ctx.array(rptr)
    S.call("senk_ilu0_f64", ctx.h, dv.p, dc.p, dr.p, N, None, 0, 1, O.P(rptr), O.P(cind))
    assert np.array_equal(dv.download(), fo)
    v32 = val.astype(np.float32)
    fo32 = v32.copy()
    assert O.orc().or_ilu0_f(O.P(fo32), O.P(cind), O.P(rptr), N) == 0
    d32 = ctx.array(v32)
    S.call("senk_ilu0_f32", ctx.h, d32.p, dc.p, dr.p, N, None, 0, 1, O.P(rptr), O.P(cind))
    assert np.array_equal(d32.download(), fo32)


def test_ilu0_zero_pivot_reported(ctx):
    # 2x2 with a structurally present zero diagonal in row 0: Ilu0 divides by it
    rptr = np.array([0, 2, 4], np.int32)
    cind = np.array([0, 1, 0, 1], np.int32)
    val = np.array([0.0, 1.0, 1.0, 1.0])
    dv, dc, dr = ctx.array(val), ctx.array(cind), ctx.array(rptr)
    st = S.lib().senk_ilu0_f64(ctx.h, dv.p, dc.p, dr.p, 2, None, 0, 1, O.P(rptr), O.P(cind))
    assert st == 3  # SENK_ERR_ZERO_PIVOT


@pytest.mark.parametrize("width", [300, 1500, 3000])
def test_ilu0_device_long_rows(ctx, width):
    """Rows longer than 256 entries (ILU(p) fill of wide rows): the factorization keeps a
    row in registers, 64 lanes x R entries, with R chosen from the longest row (4, 16,
    64).  A banded matrix whose last rows are dense over their first `width` columns, in
    FP64 and FP32, bit for bit against the oracle's Ilu0."""
    rng = np.random.default_rng(width)
    N = width + 40
    rows = []
    for i in range(N):
        cols = set(range(max(0, i - 2), min(N, i + 3)))
        if i >= N - 4:
            cols |= set(range(0, width))
        rows.append(sorted(cols))
    rptr = np.zeros(N + 1, np.int32)
    rptr[1:] = np.cumsum([len(r) for r in rows])
    cind = np.array([c for r in rows for c in r], np.int32)
    val = rng.uniform(-1, 1, len(cind))
    for i in range(N):  # diagonally dominant
        s, e = rptr[i], rptr[i + 1]
        d = s + rows[i].index(i)
        val[d] = np.sum(np.abs(val[s:e])) + 1.0
    assert np.max(np.diff(rptr)) >= width
    fo = oracle_ilu0(val, cind, rptr)
    dv, dc, dr = ctx.array(val), ctx.array(cind), ctx.array(rptr)
    S.call("senk_ilu0_f64", ctx.h, dv.p, dc.p, dr.p, N, None, 0, 1, O.P(rptr), O.P(cind))
    assert np.array_equal(dv.download(), fo)
    v32 = val.astype(np.float32)
    fo32 = v32.copy()
    assert O.orc().or_ilu0_f(O.P(fo32), O.P(cind), O.P(rptr), N) == 0
    d32 = ctx.array(v32)
    S.call("senk_ilu0_f32", ctx.h, d32.p, dc.p, dr.p, N, None, 0, 1, O.P(rptr), O.P(cind))
    assert np.array_equal(d32.download(), fo32)


def test_blas1(ctx):
    rng = np.random.default_rng(12)
    N = 100003
    x, y = rng.uniform(-1, 1, N), rng.uniform(-1, 1, N)
    dx, dy = ctx.array(x), ctx.array(y)
    a, b = 0.37, -1.25
    S.call("senk_axpy_f64", ctx.h, C.c_double(a), dx.p, dy.p, N)
    assert np.array_equal(dy.download(), y + a * x)
    dy.upload(y)
    S.call("senk_axpby_f64", ctx.h, C.c_double(a), dx.p, C.c_double(b), dy.p, N)
    assert np.array_equal(dy.download(), a * x + b * y)
    dz = ctx.zeros(N)
    S.call("senk_axpyz_f64", ctx.h, C.c_double(a), dx.p, dy.p, dz.p, N)
    yy = a * x + b * y
    assert np.array_equal(dz.download(), a * x + yy)
    S.call("senk_scal_f64", ctx.h, C.c_double(b), dz.p, N)
    assert np.array_equal(dz.download(), (a * x + yy) * b)
    S.call("senk_copy_f64", ctx.h, dx.p, dz.p, N)
    assert np.array_equal(dz.download(), x)
    dy.upload(y)
    S.call("senk_hadprod_f64", ctx.h, dx.p, dy.p, N)
    assert np.array_equal(dy.download(), y * x)
    dy.upload(y)
    S.call("senk_haddiv_f64", ctx.h, dx.p, dy.p, N)
    assert np.array_equal(dy.download(), y / x)
    r = C.c_double()
    S.call("senk_dot_f64", ctx.h, dx.p, dz.p, N, C.byref(r))
    exact = float(np.sum(np.abs(x * x)))
    assert abs(r.value - float(np.dot(x, x))) <= 1e-14 * exact
    r2 = C.c_double()
    S.call("senk_dot_f64", ctx.h, dx.p, dz.p, N, C.byref(r2))
    assert r.value == r2.value  # deterministic
    S.call("senk_nrm2_f64", ctx.h, dx.p, N, C.byref(r))
    assert abs(r.value - np.linalg.norm(x)) <= 1e-14 * np.linalg.norm(x)
    d32 = ctx.zeros(N, np.float32)
    S.call("senk_convert_f64_f32", ctx.h, dx.p, d32.p, N)
    assert np.array_equal(d32.download(), x.astype(np.float32))


@pytest.mark.parametrize("N", [1, 63, 64, 65, 1000, 257 * 1024 + 5])
def test_dot_sizes_wave_reduction(ctx, N):
    # integer-valued data: every partial sum is exact, so the DPP tree must give the
    # exact total for every size (lane layout check of the wave reduction)
    x = (np.arange(N) % 7 - 3).astype(np.float64)
    y = (np.arange(N) % 5 - 2).astype(np.float64)
    dx, dy = ctx.array(x), ctx.array(y)
    r = C.c_double()
    S.call("senk_dot_f64", ctx.h, dx.p, dy.p, N, C.byref(r))
    assert r.value == float(np.dot(x, y))
