"""Direct pins of device results that the solver tests otherwise see only through
iteration counts:

* Scaling on the device (senk_scaling_dev, senk_matrix.hpp:406-416) against
  or_scaling_d, bit for bit;
* the FP64 relative residual ||b - A x|| / ||b|| (test::RelativeResidualError,
  senk_test.hpp:11-22) -- the C ABI's senk_relative_residual_error and the true_relres
  every solve reports -- against or_relative_residual_d;
* FGMRES(m) with a fixed ILU(0) preconditioner is right-preconditioned GMRES(m) in exact
  arithmetic (Saad 1993), so its count must match the oracle's ILU-GMRES(m)
  (senk_gmres.hpp:197-270 / :391-465) within 5%;
* the unstructured mixed-precision workload (FP32 ILU(0) + inner GMRES(50) + FP64
  refinement) and the block path (ILUB 3x3 factors, FP32 + FP64 refinement + learned
  FGMRES(50)) against the oracle's FP64 solve of the same reordered system: ILU-GMRES(50)
  and ILUB-GMRES(50) to 1e-14, the FP64 forms of AbmcIluGmresm / AbmcIlubGmresm
  (senk_gmres.hpp:391-465, :679-753).
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from senk_amd import gnn as G
from tests.helpers import abmc_system, ilu_factors, oracle_ilu0, relerr

pytestmark = pytest.mark.gpu
o = O.orc()
P = O.P


def _system(rptr, cind, val, bsize, seed):
    N = len(rptr) - 1
    xs = S.gen_uniform(seed, N)
    b = np.zeros(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(xs), P(b), N)
    sy = abmc_system(rptr, cind, val, bsize)
    sy["b"], sy["xs"] = b[sy["LP"]].copy(), xs[sy["LP"]].copy()
    return sy


def _handles(ctx, sy, kind, p=0, bnl=0, bnw=0):
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, sy["N"], P(sy["rptr"]), P(sy["cind"]), P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(kind, p, bnl, bnw, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32), 0)
    M = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A, P(sy["rptr"]), P(sy["cind"]), P(sy["val"]),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts), C.byref(M))
    return A, M


def _solve(ctx, A, M, b, N, **kw):
    db, dx = ctx.array(b), ctx.zeros(N)
    res = S.senk_result()
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(S.solve_opts(**kw)), C.byref(res), None, 0)
    return res, dx.download()


def _oracle_solve(sy, spec, m, eps):
    N = sy["N"]
    x = np.zeros(N)
    conv = C.c_int()
    it = o.or_gmres_d(P(sy["val"]), P(sy["cind"]), P(sy["rptr"]), C.byref(spec), P(sy["b"]), P(x),
                      float(np.linalg.norm(sy["b"])), 20000 // m, m, N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    return it, x


@pytest.mark.parametrize("gen", ["vardiff7_10", "unstructured_3000", "elasticity_5"])
def test_scaling_dev_matches_oracle(ctx, gen):
    rptr, cind, val = {"vardiff7_10": lambda: S.gen_vardiff7(10, 2, 4.0, 1.5),
                       "unstructured_3000": lambda: S.gen_unstructured(3000, 5),
                       "elasticity_5": lambda: S.gen_elasticity3(5, 3)}[gen]()
    N = len(rptr) - 1
    d = np.zeros(N)
    o.or_get_diag_d(P(val), P(cind), P(rptr), P(d), N)
    rng = np.random.default_rng(3)
    for ld, rd in ((1 / np.sqrt(np.abs(d)), 1 / np.sqrt(np.abs(d))), (rng.uniform(0.5, 2, N), rng.uniform(0.5, 2, N))):
        want = val.copy()
        o.or_scaling_d(P(want), P(cind), P(rptr), P(ld), P(rd), N)
        dv, dc, dr, dl, drr = (ctx.array(t) for t in (val, cind, rptr, ld, rd))
        S.call("senk_scaling_dev", ctx.h, dv.p, dc.p, dr.p, dl.p, drr.p, N)
        assert np.array_equal(dv.download(), want)


def _relres_slack(sy, x):
    """Bound on how far two FP64 evaluations of ||b - A x|| / ||b|| may differ when they
    sum each row in a different order (the device splits a row over 8 lanes): every
    residual entry carries at most (row length + 1) eps_64 (|A||x| + |b|)_i of rounding."""
    N = sy["N"]
    ax = np.zeros(N)
    o.or_spmv_csr_d(P(np.abs(sy["val"])), P(sy["cind"]), P(sy["rptr"]), P(np.abs(x)), P(ax), N)
    w = int(np.diff(sy["rptr"]).max()) + 1
    return w * np.finfo(float).eps * np.linalg.norm(ax + np.abs(sy["b"])) / np.linalg.norm(sy["b"])


def test_relative_residual_matches_oracle(ctx):
    sy = _system(*S.gen_vardiff7(16, 4, 8.0, 1.5), 8, 4)
    N = sy["N"]
    dv, dc, dr = ctx.array(sy["val"]), ctx.array(sy["cind"]), ctx.array(sy["rptr"])
    db = ctx.array(sy["b"])
    rng = np.random.default_rng(9)
    for x in (np.zeros(N), rng.uniform(-1, 1, N), sy["xs"] + 1e-6 * rng.uniform(-1, 1, N)):
        want = o.or_relative_residual_d(P(sy["val"]), P(sy["cind"]), P(sy["rptr"]), P(sy["b"]), P(x), N)
        got = C.c_double()
        S.call("senk_relative_residual_error", ctx.h, dv.p, dc.p, dr.p, db.p, ctx.array(x).p, N, 0, C.byref(got))
        assert abs(got.value - want) <= 1e-13 * want + _relres_slack(sy, x), (got.value, want)
    assert got.value < 1e-3  # the last x is close to the solution: the slack term is what matters
    # the true residual a solve reports, against the oracle on the solution it returns
    A, M = _handles(ctx, sy, S.SENK_ILU0)
    try:
        for its, eps in ((5, 1e-30), (6000, 1e-10)):
            res, x = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=30, eps=eps, max_iter=its)
            want = o.or_relative_residual_d(P(sy["val"]), P(sy["cind"]), P(sy["rptr"]), P(sy["b"]), P(x), N)
            assert abs(res.true_relres - want) <= 1e-13 * want + _relres_slack(sy, x), (res.true_relres, want)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("gen,bsize,m", [("vardiff7_32", 8, 30), ("vardiff7_32", 8, 10), ("poisson7_32", 4, 10),
                                         ("convdiff27_24", 16, 30), ("unstructured_32768", 16, 5)])
def test_fgmres_with_fixed_ilu_matches_gmres_count(ctx, gen, bsize, m):
    rptr, cind, val = {"vardiff7_32": lambda: S.gen_vardiff7(32, 1, 8.0, 1.5),
                       "poisson7_32": lambda: S.gen_poisson7(32),
                       "convdiff27_24": lambda: S.gen_convdiff27(24, 10.0),
                       "unstructured_32768": lambda: S.gen_unstructured(32768, 6)}[gen]()
    sy = _system(rptr, cind, val, bsize, 1)
    N, eps = sy["N"], 1e-10
    L, U = ilu_factors(sy)
    it_o, xo = _oracle_solve(sy, O.prec_spec(1, L, U), m, eps)
    A, M = _handles(ctx, sy, S.SENK_ILU0)
    try:
        fg, xf = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_FGMRES, m=m, eps=eps, max_iter=20000)
        gm, xg = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=m, eps=eps, max_iter=20000)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
    for res in (fg, gm):
        assert res.converged == 1 and res.true_relres <= 10 * eps
        assert abs(res.iters - it_o) <= max(2, int(np.ceil(0.05 * it_o))), (res.iters, it_o)
    assert abs(fg.iters - gm.iters) <= max(1, int(np.ceil(0.05 * gm.iters))), (fg.iters, gm.iters)
    assert relerr(xf, xo) <= 1e-6 and relerr(xg, xo) <= 1e-6


def test_unstructured_mixed_ir_gmres50_against_oracle(ctx):
    """Unstructured nonsymmetric workload at 65,536 rows (1/64 of its benchmark size):
    FP32 ILU(0) factors and SpMV, inner GMRES(50), FP64 refinement to 1e-10; error
    stated against the oracle's FP64 ILU-GMRES(50) solve to 1e-14 and against x*."""
    sy = _system(*S.gen_unstructured(65536, 3), 16, 3)
    N = sy["N"]
    L, U = ilu_factors(sy)
    it_o, xo = _oracle_solve(sy, O.prec_spec(1, L, U), 50, 1e-14)
    assert o.or_relative_residual_d(P(sy["val"]), P(sy["cind"]), P(sy["rptr"]), P(sy["b"]), P(xo), N) <= 1e-13
    A, M = _handles(ctx, sy, S.SENK_ILU0)
    try:
        r64, x64 = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
        res, x = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-10, max_iter=20000, mixed=1,
                        inner_eps=1e-5, inner_max=500)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
    # the plain FP64 device path reproduces the oracle's count and solution
    assert abs(r64.iters - it_o) <= max(2, int(np.ceil(0.05 * it_o))), (r64.iters, it_o)
    assert relerr(x64, xo) <= 1e-12
    assert res.converged == 1 and res.status == 0 and res.true_relres <= 1e-10
    # forward error against the FP64 oracle solve: at most cond-amplified 1e-10 (this
    # diagonally dominant system is well conditioned)
    assert relerr(x, xo) <= 1e-8, relerr(x, xo)
    assert relerr(x, sy["xs"]) <= 1e-8


@pytest.mark.parametrize("learned", [0, 1])
def test_block_path_mixed_fgmres50_against_oracle(ctx, learned):
    """Block path at 12^3 nodes (5,184 rows): ILUB (3x3 block fill) over ABMC colors,
    FP32 factors, FP64 refinement, FGMRES(50) with and without the learned correction;
    error against the oracle's FP64 ILUB-GMRES(50) solve (BCSR factors) to 1e-14."""
    bnl = bnw = 3
    sy = _system(*S.gen_elasticity3(12, 1), 48, 1)
    N = sy["N"]
    fv, fc, fr = S.alloc_block_zero(sy["val"], sy["cind"], sy["rptr"], N, bnl, bnw)
    sp = S.split(oracle_ilu0(fv, fc, fr), fc, fr, N, "L-DU", True)
    bl, bu = S.csr2bcsr(*sp["L"], N, bnl, bnw), S.csr2bcsr(*sp["U"], N, bnl, bnw)
    it_o, xo = _oracle_solve(sy, O.prec_spec(2, bl, bu, bnl, bnw), 50, 1e-14)
    A, M = _handles(ctx, sy, S.SENK_ILUB, bnl=bnl, bnw=bnw)
    net = None
    try:
        kw = {}
        if learned:
            theta, _ = G.load()
            net = G.Gnn(ctx, theta)
            net.bind(A)
            net.attach(M)
            kw["use_learned"] = 1
        res, x = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000, mixed=1,
                        inner_eps=1e-5, inner_max=500, **kw)
        r64, x64 = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
    finally:
        if net is not None:
            S.call("senk_prec_attach_gnn", M, None)
            net.close()
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
    assert abs(r64.iters - it_o) <= max(2, int(np.ceil(0.05 * it_o))), (r64.iters, it_o)
    assert relerr(x64, xo) <= 1e-11
    assert res.converged == 1 and res.status == 0 and res.true_relres <= 1e-10
    assert relerr(x, xo) <= 1e-7, relerr(x, xo)
    assert res.fell_back in (0, 1)
