"""The mixed-precision workloads beside the headline system, at test sizes.

* Unstructured nonsymmetric matrix (lognormal row lengths, banded-random columns,
  diagonal 1.1x the row's absolute sum): FP32 values, FP32 ILU(0) factors, FP32 SpMV and
  an inner GMRES(50), FP64 iterative refinement to 1e-10; the error is reported against
  an FP64 solve of the same system to 1e-14 (plain GMRES on the device, whose own
  parity with the oracle drivers is tested elsewhere).
* Block path: the 3-dof elasticity-like system (3x3-block 7-point stencil, lognormal
  moduli) with ILUB (3x3 block fill) and with ILU(1) over ABMC colors, FP32 factors, FP64
  refinement and the learned correction inside FGMRES(50): converges to 1e-10 in FP64,
  the correction either helps or is switched off by the stagnation monitor, never
  breaking the solve.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from senk_amd import gnn as G
from tests.helpers import abmc_system, relerr

pytestmark = pytest.mark.gpu


def _system(rptr, cind, val, bsize, seed):
    N = len(rptr) - 1
    xs = S.gen_uniform(seed, N)
    b = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(xs), O.P(b), N)
    sy = abmc_system(rptr, cind, val, bsize)
    sy["b"], sy["xs"] = b[sy["LP"]].copy(), xs[sy["LP"]].copy()
    return sy


def _handles(ctx, sy, kind, p=0, bnl=0, bnw=0):
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(kind, p, bnl, bnw, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32), 0)
    M = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts), C.byref(M))
    return A, M


def _solve(ctx, A, M, b, N, **kw):
    db, dx = ctx.array(b), ctx.zeros(N)
    res = S.senk_result()
    o = S.solve_opts(**kw)
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0)
    return res, dx.download()


def test_unstructured_mixed_ir_gmres50(ctx):
    rptr, cind, val = S.gen_unstructured(65536, 3)
    sy = _system(rptr, cind, val, 16, 3)
    N = sy["N"]
    A, M = _handles(ctx, sy, S.SENK_ILU0)
    try:
        r64, x64 = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
        assert r64.converged and r64.true_relres <= 1e-13
        res, x = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-10, max_iter=20000, mixed=1,
                        inner_eps=1e-5, inner_max=500)
        assert res.converged == 1 and res.status == 0
        assert res.true_relres <= 1e-10
        # error against the FP64 solve at the level the residual allows
        assert relerr(x, x64) <= 1e-8, relerr(x, x64)
        assert relerr(x, sy["xs"]) <= 1e-8
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("kind", ["ilub", "ilu1"])
def test_block_path_mixed_learned_fgmres50(ctx, kind):
    rptr, cind, val = S.gen_elasticity3(12, 1)
    sy = _system(rptr, cind, val, 48, 1)
    N = sy["N"]
    if kind == "ilub":
        A, M = _handles(ctx, sy, S.SENK_ILUB, bnl=3, bnw=3)
    else:
        A, M = _handles(ctx, sy, S.SENK_ILUP, p=1)
    theta, _ = G.load()
    net = G.Gnn(ctx, theta)
    try:
        net.bind(A)
        net.attach(M)
        r64, x64 = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
        assert r64.converged
        plain, xp = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000, mixed=1,
                           inner_eps=1e-5, inner_max=500)
        learned, xl = _solve(ctx, A, M, sy["b"], N, solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000,
                             mixed=1, inner_eps=1e-5, inner_max=500, use_learned=1)
        for res, x in ((plain, xp), (learned, xl)):
            assert res.converged == 1 and res.status == 0
            assert res.true_relres <= 1e-10
            assert relerr(x, x64) <= 1e-7, relerr(x, x64)
        assert plain.fell_back == 0 and learned.fell_back in (0, 1)
    finally:
        S.call("senk_prec_attach_gnn", M, None)
        net.close()
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("case,solver,m", [(c, s, 30) for c in ("convdiff27_ilu0", "elasticity_ilub")
                                           for s in (S.SENK_BICGSTAB, S.SENK_FGMRES, S.SENK_GCR, S.SENK_FGCR)] +
                         [("vardiff7_ic0", S.SENK_CG, 30)])
def test_fp64_solve_with_fp32_factor_application(ctx, case, solver, m):
    """mixed = 3: the FP64 Krylov solver with its ILU / ILUB application run on the FP32
    factor copies (the residual converted to FP32 on the way in, the correction back to
    FP64).  The FP32 application differs from the FP64 one only by rounding, so the solve
    reaches the FP64 tolerance in about the FP64 solve's count, and its solution agrees
    with the FP64 solve's to the level the residual allows.  Without an FP32 copy, with the
    learned correction, with right-preconditioned GMRES (its end-of-cycle M^-1 (V y) would
    carry FP32 rounding the cycle's residual never saw: a true residual of ~1e-7 where the
    recurrence reports 1e-10) or with an unknown mode the call is refused."""
    if case == "convdiff27_ilu0":
        rptr, cind, val = S.gen_convdiff27(16, 10.0)
        sy = _system(rptr, cind, val, 16, 5)
        A, M = _handles(ctx, sy, S.SENK_ILU0)
    elif case == "vardiff7_ic0":
        rptr, cind, val = S.gen_vardiff7(16, 3, 4.0, 1.5)
        sy = _system(rptr, cind, val, 8, 3)
        A, M = _handles(ctx, sy, S.SENK_IC0)
    else:
        rptr, cind, val = S.gen_elasticity3(12, 2)
        sy = _system(rptr, cind, val, 48, 2)
        A, M = _handles(ctx, sy, S.SENK_ILUB, bnl=3, bnw=3)
    N = sy["N"]
    try:
        r64, x64 = _solve(ctx, A, M, sy["b"], N, solver=solver, m=m, eps=1e-10, max_iter=20000)
        r32, x32 = _solve(ctx, A, M, sy["b"], N, solver=solver, m=m, eps=1e-10, max_iter=20000, mixed=3)
        print(f"{case} solver {solver}: FP64 factors {r64.iters} its, FP32 factors {r32.iters} its", flush=True)
        for res in (r64, r32):
            assert res.converged == 1 and res.status == 0
            assert res.true_relres <= 1e-9
        assert r32.iters <= 1.1 * r64.iters + 3, (r32.iters, r64.iters)
        assert relerr(x32, x64) <= 1e-6, relerr(x32, x64)
        db, dx = ctx.array(sy["b"]), ctx.zeros(N)
        res = S.senk_result()
        for bad in (dict(mixed=4), dict(mixed=3, use_learned=1), dict(mixed=3, solver=S.SENK_GMRES)):
            o = S.solve_opts(**dict(dict(solver=solver, m=m, eps=1e-10, max_iter=100), **bad))
            assert S.lib().senk_solve(ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0) != 0, bad
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
    # a preconditioner without the FP32 copy
    A2 = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1, C.byref(A2))
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, 1 << S.SENK_FACTOR_F64, 0)
    M2 = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A2, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts), C.byref(M2))
    try:
        db, dx = ctx.array(sy["b"]), ctx.zeros(N)
        res = S.senk_result()
        o = S.solve_opts(solver=solver, m=m, eps=1e-10, max_iter=100, mixed=3)
        assert S.lib().senk_solve(ctx.h, A2, M2, db.p, dx.p, C.byref(o), C.byref(res), None, 0) != 0
    finally:
        S.lib().senk_prec_destroy(M2)
        S.lib().senk_mat_destroy(A2)
