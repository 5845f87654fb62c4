"""The sequential, AMC, block-Jacobi and BCSR drivers of the reference against the
oracle's restatement of the same drivers.

Bar as for the other drivers: the oracle's iteration count within 5% (at least +-2
iterations) and the FP64 true relative residual at the requested tolerance.  The
sequential forms run on dependency-level schedules of the factors; their triangular
solves give the sequential result bit for bit, so only reduction order separates the
two runs.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import oracle_ilu0

pytestmark = pytest.mark.gpu


def _close_count(gpu, ref):
    return abs(gpu - ref) <= max(2, int(np.ceil(0.05 * ref)))


def _rhs(rptr, cind, val, seed=1):
    N = len(rptr) - 1
    xs = S.gen_uniform(seed, N)
    b = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(val), O.P(cind), O.P(rptr), O.P(xs), O.P(b), N)
    return b


def _plans(ctx, L, U, N):
    lp, up = C.c_void_p(), C.c_void_p()
    S.call("senk_trsv_plan_create", ctx.h, O.P(L[2]), O.P(L[1]), N, 0, C.byref(lp))
    S.call("senk_trsv_plan_create", ctx.h, O.P(U[2]), O.P(U[1]), N, 1, C.byref(up))
    return lp, up


def _oracle(kind, val, cind, rptr, spec, b, N, eps, outer=200, m=30, max_iter=6000):
    xo = np.zeros(N)
    conv = C.c_int()
    nrm_b = float(np.linalg.norm(b))
    if kind == "gmres":
        it = O.orc().or_gmres_d(O.P(val), O.P(cind), O.P(rptr), C.byref(spec), O.P(b), O.P(xo), nrm_b, outer, m, N,
                                eps, None, 0, C.byref(conv))
    elif kind == "gcr":
        it = O.orc().or_gcr_d(O.P(val), O.P(cind), O.P(rptr), C.byref(spec), O.P(b), O.P(xo), nrm_b, outer, m, N,
                              eps, None, 0, C.byref(conv))
    else:
        it = O.orc().or_bicgstab_d(O.P(val), O.P(cind), O.P(rptr), C.byref(spec), O.P(b), O.P(xo), nrm_b, max_iter,
                                   N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    return it, nrm_b


@pytest.mark.parametrize("kind", ["gmres", "gcr", "bicgstab"])
def test_sequential_ilu_drivers(ctx, kind):
    # natural ordering: IluGmresm / IluGcrm / IluBicgstab with the sequential solves
    rptr, cind, val = S.gen_convdiff27(12, 10.0)
    N, eps = len(rptr) - 1, 1e-10
    b = _rhs(rptr, cind, val)
    f = oracle_ilu0(val, cind, rptr)
    sp = S.split(f, cind, rptr, N, "L-DU", True)
    L, U = sp["L"], sp["U"]
    it_o, nrm_b = _oracle(kind, val, cind, rptr, O.prec_spec(1, L, U), b, N, eps)
    lp, up = _plans(ctx, L, U, N)
    dv, dc, dr, db = ctx.array(val), ctx.array(cind), ctx.array(rptr), ctx.array(b)
    dl, du = [ctx.array(t) for t in L], [ctx.array(t) for t in U]
    dx = ctx.zeros(N)
    res = S.senk_result()
    if kind == "gmres":
        S.call("senk_ilu_gmresm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), lp, up, db.p,
               dx.p, C.c_double(nrm_b), 200, 30, N, C.c_double(eps), 0, C.byref(res), None, 0)
    elif kind == "gcr":
        S.call("senk_ilu_gcrm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), lp, up, db.p,
               dx.p, C.c_double(nrm_b), 200, 30, N, C.c_double(eps), 0, C.byref(res), None, 0)
    else:
        S.call("senk_ilu_bicgstab_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), lp, up,
               db.p, dx.p, C.c_double(nrm_b), 6000, N, C.c_double(eps), 0, C.byref(res), None, 0)
    S.lib().senk_trsv_plan_destroy(lp)
    S.lib().senk_trsv_plan_destroy(up)
    assert res.converged == 1
    if kind == "gcr":
        assert res.iters == it_o  # restart granularity
    else:
        assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= (100 if kind == "bicgstab" else 10) * eps


@pytest.mark.parametrize("kind", ["gmres", "gcr", "bicgstab"])
def test_sequential_ilub_drivers(ctx, kind):
    # IlubGmresm / IlubGcrm / IlubBicgstab: CSR matrix, BCSR ILUB factors (2x2 blocks)
    bnl = bnw = 2
    rptr, cind, val = S.gen_convdiff27(10, 10.0)
    N, eps = len(rptr) - 1, 1e-10
    b = _rhs(rptr, cind, val)
    fv, fc, fr = S.alloc_block_zero(val, cind, rptr, N, bnl, bnw)
    f = oracle_ilu0(fv, fc, fr)
    sp = S.split(f, fc, fr, N, "L-DU", True)
    bl = S.csr2bcsr(*sp["L"], N, bnl, bnw)
    bu = S.csr2bcsr(*sp["U"], N, bnl, bnw)
    it_o, nrm_b = _oracle(kind, val, cind, rptr, O.prec_spec(2, bl, bu, bnl, bnw), b, N, eps)
    dv, dc, dr, db = ctx.array(val), ctx.array(cind), ctx.array(rptr), ctx.array(b)
    dl, du = [ctx.array(t) for t in bl], [ctx.array(t) for t in bu]
    dx = ctx.zeros(N)
    res = S.senk_result()
    if kind == "gmres":
        S.call("senk_ilub_gmresm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), db.p, dx.p,
               C.c_double(nrm_b), 200, 30, N, C.c_double(eps), bnl, bnw, 0, C.byref(res), None, 0)
    elif kind == "gcr":
        S.call("senk_ilub_gcrm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), db.p, dx.p,
               C.c_double(nrm_b), 200, 30, N, C.c_double(eps), bnl, bnw, 0, C.byref(res), None, 0)
    else:
        S.call("senk_ilub_bicgstab_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), db.p,
               dx.p, C.c_double(nrm_b), 6000, N, C.c_double(eps), bnl, bnw, 0, C.byref(res), None, 0)
    assert res.converged == 1
    if kind == "gcr":
        assert res.iters == it_o
    else:
        assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= (100 if kind == "bicgstab" else 10) * eps


def test_bj_ilu_gmresm(ctx):
    # BjIluGmresm: ILU(0) of the block-diagonal part (bnum ranges), sequential solves per range
    bnum = 8
    rptr, cind, val = S.gen_poisson7(12)
    N, eps = len(rptr) - 1, 1e-8
    b = _rhs(rptr, cind, val)
    rv, rc, rr = S.remove_off_diagonal(val, cind, rptr, N, bnum)
    f = oracle_ilu0(rv, rc, rr)
    sp = S.split(f, rc, rr, N, "L-DU", True)
    L, U = sp["L"], sp["U"]
    it_o, nrm_b = _oracle("gmres", val, cind, rptr, O.prec_spec(1, L, U), b, N, eps)
    lp, up = _plans(ctx, L, U, N)
    dv, dc, dr, db = ctx.array(val), ctx.array(cind), ctx.array(rptr), ctx.array(b)
    dl, du = [ctx.array(t) for t in L], [ctx.array(t) for t in U]
    dx = ctx.zeros(N)
    res = S.senk_result()
    S.call("senk_bj_ilu_gmresm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), lp, up, bnum,
           db.p, dx.p, C.c_double(nrm_b), 200, 30, N, C.c_double(eps), 0, C.byref(res), None, 0)
    S.lib().senk_trsv_plan_destroy(lp)
    S.lib().senk_trsv_plan_destroy(up)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= 10 * eps


def test_amc_ilu_gmresm(ctx):
    rptr, cind, val = S.gen_convdiff27(10, 10.0)
    N, eps = len(rptr) - 1, 1e-10
    nc, sc, LP, RP = S.amc_permutation(cind, rptr, N)
    v2, c2, r2 = S.reordering(val, cind, rptr, LP, RP, N)
    b = _rhs(r2, c2, v2)
    f = oracle_ilu0(v2, c2, r2)
    sp = S.split(f, c2, r2, N, "L-DU", True)
    L, U = sp["L"], sp["U"]
    it_o, nrm_b = _oracle("gmres", v2, c2, r2, O.prec_spec(1, L, U), b, N, eps)
    dv, dc, dr, db = ctx.array(v2), ctx.array(c2), ctx.array(r2), ctx.array(b)
    dl, du = [ctx.array(t) for t in L], [ctx.array(t) for t in U]
    dx = ctx.zeros(N)
    res = S.senk_result()
    cp = sc.ctypes.data_as(C.c_void_p)
    S.call("senk_amc_ilu_gmresm_f64", ctx.h, dv.p, dc.p, dr.p, *(t.p for t in dl), *(t.p for t in du), cp, nc, db.p,
           dx.p, C.c_double(nrm_b), 200, 30, N, C.c_double(eps), 0, C.byref(res), None, 0)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    assert res.true_relres <= 10 * eps


def test_gmresm_bcsr_unpreconditioned(ctx):
    # Gmresm on a BCSR matrix against the oracle's CSR GMRES on the same matrix
    # (BCSR -> CSR keeps the explicit zeros of the blocks; same operator)
    bnl = bnw = 2
    rptr, cind, val = S.gen_convdiff27(8, 10.0)
    N, eps = len(rptr) - 1, 1e-8
    bv, bc, br = S.csr2bcsr(val, cind, rptr, N, bnl, bnw)
    v2, c2, r2 = S.bcsr2csr(bv, bc, br, N, bnl, bnw)
    b = _rhs(rptr, cind, val)
    it_o, nrm_b = _oracle("gmres", v2, c2, r2, O.prec_spec(0), b, N, eps, outer=400)
    dv, dc, dr, db = ctx.array(bv), ctx.array(bc), ctx.array(br), ctx.array(b)
    dx = ctx.zeros(N)
    res = S.senk_result()
    S.call("senk_gmresm_bcsr_f64", ctx.h, dv.p, dc.p, dr.p, db.p, dx.p, C.c_double(nrm_b), 400, 30, N,
           C.c_double(eps), bnl, bnw, 0, C.byref(res), None, 0)
    assert res.converged == 1
    assert _close_count(res.iters, it_o), (res.iters, it_o)
    # the FP64 true residual is recomputed with the BCSR SpMV
    assert res.true_relres <= 10 * eps
