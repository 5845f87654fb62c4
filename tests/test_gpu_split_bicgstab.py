"""Split (Eisenstat) form of the mixed forms' inner FP32 BiCGStab, on the GPU.

For an ILU(0) that changes no off-diagonal entry (checked on the pattern when the
preconditioner is created; the algebra is pinned on the CPU in test_split_identity.py)
the inner BiCGStab runs on L^-1 A U^-1 with one U and one L solve per product and no
SpMV.  Checks: the nested and refinement solves converge to the same FP64 tolerance and
error against x* with the split form on and off, in every triangular-solve launch mode,
with iteration counts close to each other; the single-triangle launches keep the
sweeps' state, so plain applications before and after split solves agree bit for bit
and a repeated solve repeats its iterations; where the form does not apply (27-point
stencil, FP16 factors) switching it on is refused.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, relerr

pytestmark = pytest.mark.gpu

BITS = (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32)


def _handles(ctx, sy, bits=BITS):
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, bits, 0)
    M = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts), C.byref(M))
    return A, M


def _system(n=24, bsize=8):
    rptr, cind, val = S.gen_vardiff7(n, 3, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, bsize)
    N = sy["N"]
    xs = np.random.default_rng(4).uniform(-1, 1, N)
    rows = np.repeat(np.arange(N), np.diff(sy["rptr"]))
    sy["b"] = np.bincount(rows, weights=sy["val"] * xs[sy["cind"]], minlength=N)
    sy["xs"] = xs
    return sy


def _solve(ctx, A, M, sy, mixed):
    db, dx = ctx.array(sy["b"]), ctx.zeros(sy["N"])
    o = S.solve_opts(solver=S.SENK_BICGSTAB, m=30, eps=1e-10, mixed=mixed, inner_eps=1e-3, inner_max=40,
                     outer_m=20, max_iter=6000)
    res = S.senk_result()
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0)
    return res, dx.download()


@pytest.mark.parametrize("mode", [2, 3, 1, 0])
def test_split_and_plain_inner_solves_agree(ctx, mode):
    sy = _system()
    A, M = _handles(ctx, sy)
    try:
        S.call("senk_prec_set_trsv_mode", M, mode)
        its = {}
        for mixed in (2, 1):
            for split in (1, 0):
                S.call("senk_prec_set_split", M, split)
                res, x = _solve(ctx, A, M, sy, mixed)
                assert res.converged == 1 and res.status == 0, (mixed, split)
                assert res.true_relres <= 1.5e-10, (mixed, split, res.true_relres)
                assert relerr(x, sy["xs"]) <= 1e-5, (mixed, split)
                its[(mixed, split)] = res.iters
            assert its[(mixed, 1)] <= 1.3 * its[(mixed, 0)] + 10, its
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_split_keeps_the_sweep_state(ctx):
    sy = _system(16, 4)
    N = sy["N"]
    A, M = _handles(ctx, sy)
    try:
        x = np.random.default_rng(5).uniform(-1, 1, N).astype(np.float32)
        dx, dy = ctx.array(x), ctx.zeros(N, np.float32)
        S.call("senk_prec_apply_f32", M, dx.p, dy.p)
        z0 = dy.download()
        S.call("senk_prec_set_split", M, 1)
        r1, x1 = _solve(ctx, A, M, sy, 2)
        S.call("senk_prec_apply_f32", M, dx.p, dy.p)
        assert np.array_equal(dy.download(), z0)
        r2, x2 = _solve(ctx, A, M, sy, 2)
        assert r1.converged and r2.converged
        assert r1.iters == r2.iters and np.array_equal(x1, x2)
        S.call("senk_prec_apply_f32", M, dx.p, dy.p)
        assert np.array_equal(dy.download(), z0)
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_split_refused_where_it_does_not_apply(ctx):
    rptr, cind, val = S.gen_convdiff27(8, 10.0)
    sy = abmc_system(rptr, cind, val, 8)
    A, M = _handles(ctx, sy)
    try:
        assert S.lib().senk_prec_set_split(M, 1) == 9  # SENK_ERR_UNSUPPORTED
        assert S.lib().senk_prec_set_split(M, 0) == 0
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
    sy = _system(8, 4)
    A, M = _handles(ctx, sy, BITS | (1 << S.SENK_FACTOR_F16))
    try:
        assert S.lib().senk_prec_set_split(M, 1) == 9
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
