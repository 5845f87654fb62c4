"""GPU Krylov drivers against the oracle's restatement of the reference drivers.

Bar (from the plain FP64 path's parity contract): the iteration count the reference
reports must be reproduced within 5% (at least +-2 iterations), and the FP64 true
relative residual ||b - A x|| / ||b|| must reach the requested tolerance (the stop tests
are on the implicit / recursive residual, so a small factor of slack is allowed, as for
the reference itself).  Reductions are not bit-reproducible between the OpenMP
reference and any other implementation, so counts are compared, not bits.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, ilu_factors, oracle_ilu0, relerr

pytestmark = pytest.mark.gpu


def _problem(kind, n, bsize, seed=1):
    if kind == "poisson":
        rptr, cind, val = S.gen_poisson7(n)
    elif kind == "convdiff":
        rptr, cind, val = S.gen_convdiff27(n, 10.0)
    elif kind == "vardiff":
        rptr, cind, val = S.gen_vardiff7(n, seed, 8.0, 1.5)
    elif kind == "elasticity":
        rptr, cind, val = S.gen_elasticity3(n, seed)
    else:
        raise ValueError(kind)
    N = len(rptr) - 1
    xs = S.gen_uniform(seed, N)
    b = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(xs), O.P(b), N)
    sy = abmc_system(rptr, cind, val, bsize)
    sy["b"] = b[sy["LP"]].copy()  # b permuted with the rows
    sy["xs"] = xs[sy["LP"]].copy()
    return sy


def _close_count(gpu, ref):
    # 5% of the reference count; at small counts BiCGStab's irregular residual curve can
    # cross the tolerance a step or two apart under a different reduction order, so at
    # least +-2 iterations are allowed.
    return abs(gpu - ref) <= max(2, int(np.ceil(0.05 * ref)))


def _upload_system(ctx, sy, L, U):
    d = {k: ctx.array(sy[k]) for k in ("val", "cind", "rptr", "b")}
    d["L"] = [ctx.array(t) for t in L]
    d["U"] = [ctx.array(t) for t in U]
    d["x"] = ctx.zeros(sy["N"])
    return d


@pytest.mark.parametrize("kind,n,bsize,m,eps", [("poisson", 32, 4, 30, 1e-8), ("convdiff", 16, 16, 30, 1e-10),
                                                ("vardiff", 16, 8, 30, 1e-10)])
def test_abmc_ilu_gmresm(ctx, kind, n, bsize, m, eps):
    sy = _problem(kind, n, bsize)
    N = sy["N"]
    L, U = ilu_factors(sy)
    nrm_b = float(np.linalg.norm(sy["b"]))
    outer = 6000 // m
    xo = np.zeros(N)
    conv = C.c_int()
    it_o = O.orc().or_gmres_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(1, L, U)),
                              O.P(sy["b"]), O.P(xo), nrm_b, outer, m, N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    d = _upload_system(ctx, sy, L, U)
    res = S.senk_result()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_abmc_ilu_gmresm_f64", ctx.h, d["val"].p, d["cind"].p, d["rptr"].p, *(t.p for t in d["L"]),
           *(t.p for t in d["U"]), cp, sy["nc"], bsize, d["b"].p, d["x"].p, C.c_double(nrm_b), outer, m, N,
           C.c_double(eps), 0, C.byref(res), None, 0)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= 10 * eps
    x = d["x"].download()
    assert relerr(x, xo) <= 1e3 * eps * 10


def test_gmres_cgs2_matches_mgs_count(ctx):
    sy = _problem("convdiff", 16, 16)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1, C.byref(A))
    M = C.c_void_p()
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32))
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), cp, sy["nc"], 16,
           C.byref(opts), C.byref(M))
    db = ctx.array(sy["b"])
    its = {}
    for ortho in (S.SENK_MGS, S.SENK_CGS2):
        dx = ctx.zeros(N)
        o = S.solve_opts(solver=S.SENK_GMRES, m=30, eps=1e-10, ortho=ortho, max_iter=3000)
        res = S.senk_result()
        S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0)
        assert res.converged and res.true_relres <= 1e-9
        its[ortho] = res.iters
    assert abs(its[0] - its[1]) <= max(1, int(0.05 * its[0])), its
    S.lib().senk_prec_destroy(M)
    S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("kind,n,bsize", [("convdiff", 16, 16), ("poisson", 20, 4)])
def test_abmc_ilu_bicgstab(ctx, kind, n, bsize):
    sy = _problem(kind, n, bsize)
    N, eps = sy["N"], 1e-10
    L, U = ilu_factors(sy)
    nrm_b = float(np.linalg.norm(sy["b"]))
    xo = np.zeros(N)
    conv = C.c_int()
    it_o = O.orc().or_bicgstab_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(1, L, U)),
                                 O.P(sy["b"]), O.P(xo), nrm_b, 6000, N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    d = _upload_system(ctx, sy, L, U)
    res = S.senk_result()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_abmc_ilu_bicgstab_f64", ctx.h, d["val"].p, d["cind"].p, d["rptr"].p, *(t.p for t in d["L"]),
           *(t.p for t in d["U"]), cp, sy["nc"], bsize, d["b"].p, d["x"].p, C.c_double(nrm_b), 6000, N,
           C.c_double(eps), 0, C.byref(res), None, 0)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= 100 * eps


def test_abmc_ilu_gcrm(ctx):
    sy = _problem("poisson", 16, 8)
    N, eps, m = sy["N"], 1e-10, 30
    L, U = ilu_factors(sy)
    nrm_b = float(np.linalg.norm(sy["b"]))
    xo = np.zeros(N)
    conv = C.c_int()
    ho = np.zeros(64)
    it_o = O.orc().or_gcr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(1, L, U)),
                            O.P(sy["b"]), O.P(xo), nrm_b, 200, m, N, eps, O.P(ho), 64, C.byref(conv))
    assert conv.value == 1
    d = _upload_system(ctx, sy, L, U)
    res = S.senk_result()
    hist = np.zeros(64)
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_abmc_ilu_gcrm_f64", ctx.h, d["val"].p, d["cind"].p, d["rptr"].p, *(t.p for t in d["L"]),
           *(t.p for t in d["U"]), cp, sy["nc"], 8, d["b"].p, d["x"].p, C.c_double(nrm_b), 200, m, N,
           C.c_double(eps), 0, C.byref(res), O.P(hist), 64)
    assert res.converged == 1
    assert res.iters == it_o  # restart granularity: i*m of the last restart
    assert res.true_relres <= eps
    k = res.restarts
    assert relerr(hist[:k], ho[:k]) <= 1e-6


def test_unpreconditioned_drivers(ctx):
    rptr, cind, val = S.gen_poisson7(10)
    N = len(rptr) - 1
    b = S.gen_uniform(3, N)
    nrm_b = float(np.linalg.norm(b))
    dv, dc, dr, db = ctx.array(val), ctx.array(cind), ctx.array(rptr), ctx.array(b)
    none = O.prec_spec(0)
    for name, it_fn in (("gmres", "senk_gmresm_f64"), ("gcr", "senk_gcrm_f64"), ("bicgstab", "senk_bicgstab_f64")):
        xo = np.zeros(N)
        conv = C.c_int()
        ho, hg = np.zeros(4096), np.zeros(4096)
        if name == "gmres":
            it_o = O.orc().or_gmres_d(O.P(val), O.P(cind), O.P(rptr), C.byref(none), O.P(b), O.P(xo), nrm_b, 100, 30,
                                      N, 1e-8, None, 0, C.byref(conv))
        elif name == "gcr":
            it_o = O.orc().or_gcr_d(O.P(val), O.P(cind), O.P(rptr), C.byref(none), O.P(b), O.P(xo), nrm_b, 100, 30, N,
                                    1e-8, None, 0, C.byref(conv))
        else:
            it_o = O.orc().or_bicgstab_d(O.P(val), O.P(cind), O.P(rptr), C.byref(none), O.P(b), O.P(xo), nrm_b, 3000,
                                         N, 1e-8, O.P(ho), 4096, C.byref(conv))
        dx = ctx.zeros(N)
        res = S.senk_result()
        if name == "bicgstab":
            S.call(it_fn, ctx.h, dv.p, dc.p, dr.p, db.p, dx.p, C.c_double(nrm_b), 3000, N, C.c_double(1e-8), 0,
                   C.byref(res), O.P(hg), 4096)
            # same recurrence: the residual histories agree to rounding for the first steps
            assert np.max(np.abs(hg[:10] / ho[:10] - 1)) <= 1e-8, (hg[:10], ho[:10])
        else:
            S.call(it_fn, ctx.h, dv.p, dc.p, dr.p, db.p, dx.p, C.c_double(nrm_b), 100, 30, N, C.c_double(1e-8), 0,
                   C.byref(res), None, 0)
        assert res.converged == 1, name
        assert _close_count(res.iters, it_o), (name, res.iters, it_o)
        assert res.true_relres <= 1e-7, name


def test_abmc_ilub_gmresm_bcsr(ctx):
    bnl = bnw = 3
    bsize = 12
    sy = _problem("elasticity", 6, bsize)
    N, eps, m = sy["N"], 1e-10, 50
    fv, fc, fr = S.alloc_block_zero(sy["val"], sy["cind"], sy["rptr"], N, bnl, bnw)
    f = oracle_ilu0(fv, fc, fr)
    sp = S.split(f, fc, fr, N, "L-DU", True)
    bl = S.csr2bcsr(*sp["L"], N, bnl, bnw)
    bu = S.csr2bcsr(*sp["U"], N, bnl, bnw)
    nrm_b = float(np.linalg.norm(sy["b"]))
    xo = np.zeros(N)
    conv = C.c_int()
    it_o = O.orc().or_gmres_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(2, bl, bu, bnl, bnw)),
                              O.P(sy["b"]), O.P(xo), nrm_b, 100, m, N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    d = {k: ctx.array(sy[k]) for k in ("val", "cind", "rptr", "b")}
    dl = [ctx.array(t) for t in bl]
    du = [ctx.array(t) for t in bu]
    dx = ctx.zeros(N)
    res = S.senk_result()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_abmc_ilub_gmresm_f64", ctx.h, d["val"].p, d["cind"].p, d["rptr"].p, *(t.p for t in dl),
           *(t.p for t in du), cp, sy["nc"], bsize, d["b"].p, dx.p, C.c_double(nrm_b), 100, m, N, C.c_double(eps),
           bnl, bnw, 0, C.byref(res), None, 0)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= 10 * eps


@pytest.mark.parametrize("kind,p", [(S.SENK_ILU0, 0), (S.SENK_ILUP, 1), (S.SENK_ILUB, 0)])
def test_handle_ilu_factors_match_oracle(ctx, kind, p):
    sy = _problem("convdiff", 10, 8)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 0, C.byref(A))
    opts = S.senk_ilu_opts(kind, p, 2, 2, 1 << S.SENK_FACTOR_F64)
    M = C.c_void_p()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), cp, sy["nc"], 8,
           C.byref(opts), C.byref(M))
    # oracle: symbolic pattern, Ilu0, Split L-DU
    if kind == S.SENK_ILUP:
        fv, fc, fr = S.alloc_level_zero(sy["val"], sy["cind"], sy["rptr"], N, p)
    elif kind == S.SENK_ILUB:
        fv, fc, fr = S.alloc_block_zero(sy["val"], sy["cind"], sy["rptr"], N, 2, 2)
    else:
        fv, fc, fr = sy["val"], sy["cind"], sy["rptr"]
    f = oracle_ilu0(fv, fc, fr)
    sp = S.split(f, fc, fr, N, "L-DU", True)
    ln, un, nl, nu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    S.call("senk_prec_factor_nnz", M, C.byref(ln), C.byref(un), C.byref(nl), C.byref(nu))
    assert ln.value == len(sp["L"][0]) and un.value == len(sp["U"][0])
    lr, lc, lv = np.zeros(N + 1, np.int32), np.zeros(ln.value, np.int32), np.zeros(ln.value)
    ur, uc, uv = np.zeros(N + 1, np.int32), np.zeros(un.value, np.int32), np.zeros(un.value)
    S.call("senk_prec_download", M, O.P(lr), O.P(lc), O.P(lv), O.P(ur), O.P(uc), O.P(uv))
    assert np.array_equal(lr, sp["L"][2]) and np.array_equal(lc, sp["L"][1]) and np.array_equal(lv, sp["L"][0])
    assert np.array_equal(ur, sp["U"][2]) and np.array_equal(uc, sp["U"][1]) and np.array_equal(uv, sp["U"][0])
    # apply == oracle sequential solve
    x = np.random.default_rng(4).uniform(-1, 1, N)
    yl, yu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_l_d(O.P(sp["L"][0]), O.P(sp["L"][1]), O.P(sp["L"][2]), O.P(x), O.P(yl), N)
    O.orc().or_sptrsv_u_d(O.P(sp["U"][0]), O.P(sp["U"][1]), O.P(sp["U"][2]), O.P(yl), O.P(yu), N)
    dx, dy = ctx.array(x), ctx.zeros(N)
    S.call("senk_prec_apply_f64", M, dx.p, dy.p)
    assert np.array_equal(dy.download(), yu)
    S.lib().senk_prec_destroy(M)
    S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("solver", [S.SENK_GMRES, S.SENK_FGMRES, S.SENK_GCR, S.SENK_FGCR, S.SENK_BICGSTAB])
def test_mixed_precision_ir(ctx, solver):
    sy = _problem("vardiff", 20, 16, seed=2)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32))
    M = C.c_void_p()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), cp, sy["nc"], 16,
           C.byref(opts), C.byref(M))
    db = ctx.array(sy["b"])
    # FP64 oracle-grade reference solve (plain FP64 GMRES to 1e-14)
    dx64 = ctx.zeros(N)
    o64 = S.solve_opts(solver=S.SENK_GMRES, m=30, eps=1e-14, max_iter=6000)
    r64 = S.senk_result()
    S.call("senk_solve", ctx.h, A, M, db.p, dx64.p, C.byref(o64), C.byref(r64), None, 0)
    assert r64.true_relres <= 1e-13
    dx = ctx.zeros(N)
    o = S.solve_opts(solver=solver, m=30, eps=1e-10, mixed=1, inner_eps=1e-5, inner_max=300, max_iter=6000)
    res = S.senk_result()
    hist = np.zeros(64)
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), O.P(hist), 64)
    assert res.converged == 1
    assert res.true_relres <= 1e-10
    # FP64 refinement residuals decrease monotonically
    h = hist[: res.restarts + 1]
    assert np.all(np.diff(h) < 0), h
    x, x64 = dx.download(), dx64.download()
    # error against the FP64 solve: bounded by cond(A) * 1e-10 in norm; check relative
    assert relerr(x, x64) <= 1e-7
    assert relerr(x, sy["xs"]) <= 1e-6
    S.lib().senk_prec_destroy(M)
    S.lib().senk_mat_destroy(A)


def test_reference_lines_printed(ctx, capfd):
    sy = _problem("poisson", 12, 4)
    N = sy["N"]
    L, U = ilu_factors(sy)
    d = _upload_system(ctx, sy, L, U)
    nrm_b = float(np.linalg.norm(sy["b"]))
    res = S.senk_result()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_abmc_ilu_gmresm_f64", ctx.h, d["val"].p, d["cind"].p, d["rptr"].p, *(t.p for t in d["L"]),
           *(t.p for t in d["U"]), cp, sy["nc"], 4, d["b"].p, d["x"].p, C.c_double(nrm_b), 100, 30, N,
           C.c_double(1e-8), 1, C.byref(res), None, 0)
    out = capfd.readouterr().out
    assert f"I iter {res.iters}\n" in out
    assert "# res " in out


@pytest.mark.parametrize("solver", [S.SENK_GMRES, S.SENK_FGCR])
def test_nested_mixed_precision(ctx, solver):
    # FP64 flexible GCR outer iteration around FP32 inner Krylov solves
    sy = _problem("vardiff", 20, 16, seed=3)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32))
    M = C.c_void_p()
    cp = sy["cptr"].ctypes.data_as(C.c_void_p)
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), cp, sy["nc"], 16,
           C.byref(opts), C.byref(M))
    db = ctx.array(sy["b"])
    dx64 = ctx.zeros(N)
    r64 = S.senk_result()
    o64 = S.solve_opts(solver=S.SENK_GMRES, m=30, eps=1e-14, max_iter=6000)
    S.call("senk_solve", ctx.h, A, M, db.p, dx64.p, C.byref(o64), C.byref(r64), None, 0)
    dx = ctx.zeros(N)
    o = S.solve_opts(solver=solver, m=30, eps=1e-10, mixed=2, inner_eps=1e-3, inner_max=30, outer_m=20,
                     max_iter=6000)
    res = S.senk_result()
    hist = np.zeros(256)
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), O.P(hist), 256)
    assert res.converged == 1 and res.status == 0
    assert res.true_relres <= 1e-10
    assert res.inner_iters == res.iters > 0
    # each outer GCR step minimises ||r|| over the new direction, so the FP64 residual
    # does not grow (recorded: start of each cycle, then every outer step; a restart
    # records the true residual, which differs from the recursive one by rounding of
    # order 1e-16 ||b||, i.e. up to ~1e-6 relative near 1e-10 ||b||)
    h = hist[hist > 0]
    assert len(h) >= res.restarts
    assert np.all(h[1:] <= h[:-1] * (1 + 1e-4)), h
    assert relerr(dx.download(), dx64.download()) <= 1e-7
    S.lib().senk_prec_destroy(M)
    S.lib().senk_mat_destroy(A)
