"""The sliced solve layout of the ILU factors against the CSR kernels and the oracle.

A preconditioner handle keeps its factors in the sliced layout by default and in CSR
only when asked (csr_only = 1).  Both paths must give the oracle's sequential
L-then-U solve bit for bit in FP64 (same operations in the same order per row), and
the same bits as each other in FP32, over every block size the layout takes, for
color schedules and for dependency-level schedules (ILU(p) fill, natural ordering).
The layout solve runs in each launch mode (one launch per level; one launch per
triangle whose gathers wait for the rows they read, cooperative or with workgroup
tickets): all give the same bits, also when applied many times in a row (the sweep's
intermediate vector must return to the sentinel and the ticket count stay in step with
the host after every application) and in place.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, oracle_ilu0

pytestmark = pytest.mark.gpu
MODES = (0, 1, 2, 3)  # per-level launches, cooperative sweep, ticketed sweep, per-XCD ticket streams


def _expected(sy, kind, p):
    N = sy["N"]
    if kind == S.SENK_ILUP and p > 0:
        fv, fc, fr = S.alloc_level_zero(sy["val"], sy["cind"], sy["rptr"], N, p)
    else:
        fv, fc, fr = sy["val"], sy["cind"], sy["rptr"]
    f = oracle_ilu0(fv, fc, fr)
    return S.split(f, fc, fr, N, "L-DU", True)


def _oracle_apply(sp, x):
    N = len(x)
    yl, yu = np.zeros(N), np.zeros(N)
    O.orc().or_sptrsv_l_d(O.P(sp["L"][0]), O.P(sp["L"][1]), O.P(sp["L"][2]), O.P(x), O.P(yl), N)
    O.orc().or_sptrsv_u_d(O.P(sp["U"][0]), O.P(sp["U"][1]), O.P(sp["U"][2]), O.P(yl), O.P(yu), N)
    return yu


def _handles(ctx, sy, kind, p, bsize, colors):
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 0, C.byref(A))
    out = []
    for csr_only in (0, 1):
        opts = S.senk_ilu_opts(kind, p, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32), csr_only)
        M = C.c_void_p()
        cp = sy["cptr"].ctypes.data_as(C.c_void_p) if colors else None
        S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), cp,
               sy["nc"] if colors else 0, bsize, C.byref(opts), C.byref(M))
        out.append(M)
    return A, out


def _check(ctx, sy, kind, p, bsize, colors, seed=3):
    N = sy["N"]
    A, (Ml, Mc) = _handles(ctx, sy, kind, p, bsize, colors)
    try:
        sp = _expected(sy, kind, p)
        x = np.random.default_rng(seed).uniform(-1, 1, N)
        want = _oracle_apply(sp, x)
        dx = ctx.array(x)
        dy = ctx.zeros(N)
        S.call("senk_prec_apply_f64", Mc, dx.p, dy.p)
        assert np.array_equal(dy.download(), want)
        x32 = x.astype(np.float32)
        dx32 = ctx.array(x32)
        dy32 = ctx.array(np.zeros(N, np.float32))
        S.call("senk_prec_apply_f32", Mc, dx32.p, dy32.p)
        want32 = dy32.download()
        assert np.max(np.abs(want32 - want)) <= 1e-4 * max(1.0, np.max(np.abs(want)))
        for mode in MODES:
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            for rep in range(2):  # a second application starts from the counters the first left
                dy = ctx.zeros(N)
                S.call("senk_prec_apply_f64", Ml, dx.p, dy.p)
                got = dy.download()
                assert np.array_equal(got, want), (mode, rep, np.max(np.abs(got - want)))
                dy32 = ctx.array(np.zeros(N, np.float32))
                S.call("senk_prec_apply_f32", Ml, dx32.p, dy32.p)
                assert np.array_equal(dy32.download(), want32), (mode, rep)
            S.call("senk_prec_check", Ml)
    finally:
        for M in (Ml, Mc):
            S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("bsize", [1, 2, 4, 8, 16, 32, 64])
def test_layout_abmc_poisson(ctx, bsize):
    rptr, cind, val = S.gen_poisson7(16)
    sy = abmc_system(rptr, cind, val, bsize)
    _check(ctx, sy, S.SENK_ILU0, 0, bsize, colors=True)


@pytest.mark.parametrize("bsize", [6, 12, 24, 48])
def test_layout_blocks_not_dividing_64(ctx, bsize):
    """ABMC blocks that a wave of 64 lanes cannot hold whole are laid out as sub-blocks of
    gcd(bsize, 64) rows, each color split into bsize / gcd levels in sub-block order; every
    launch mode gives the oracle's sequential solve bit for bit (12^3 = 1728 rows, a multiple
    of each block size).  The CSR path for such blocks (one wave per block, the row's
    products summed across the wave in double) is not order-preserving: it agrees to
    rounding only."""
    rptr, cind, val = S.gen_poisson7(12)
    sy = abmc_system(rptr, cind, val, bsize)
    N = sy["N"]
    A, (Ml, Mc) = _handles(ctx, sy, S.SENK_ILU0, 0, bsize, True)
    try:
        x = np.random.default_rng(3).uniform(-1, 1, N)
        want = _oracle_apply(_expected(sy, S.SENK_ILU0, 0), x)
        dx = ctx.array(x)
        dy = ctx.zeros(N)
        S.call("senk_prec_apply_f64", Mc, dx.p, dy.p)
        assert np.max(np.abs(dy.download() - want)) <= 1e-13 * np.max(np.abs(want))
        for mode in MODES:
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            for rep in range(2):
                dy = ctx.zeros(N)
                S.call("senk_prec_apply_f64", Ml, dx.p, dy.p)
                got = dy.download()
                assert np.array_equal(got, want), (mode, rep, np.max(np.abs(got - want)))
            S.call("senk_prec_check", Ml)
    finally:
        for M in (Ml, Mc):
            S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("bsize", [4, 16])
def test_layout_abmc_convdiff27(ctx, bsize):
    rptr, cind, val = S.gen_convdiff27(12, 10.0)
    sy = abmc_system(rptr, cind, val, bsize)
    _check(ctx, sy, S.SENK_ILU0, 0, bsize, colors=True)


def test_layout_abmc_unstructured_long_rows(ctx):
    # lognormal row lengths: rows of up to 37 (L) and 39 (U) entries outside their block,
    # so the sweeps take their gathers in pipelined rounds of 16 (more than 32) -- and of 4
    # for FP64 in the other tests' short rows, of 8 with pipelining in the 27-point ones
    rptr, cind, val = S.gen_unstructured(16384, 3)
    sy = abmc_system(rptr, cind, val, 16)
    rp, ci = sy["rptr"], sy["cind"]
    ext = [int(np.sum((ci[rp[i]:rp[i + 1]] < (i // 16) * 16) | (ci[rp[i]:rp[i + 1]] >= (i // 16) * 16 + 16)))
           for i in range(sy["N"])]
    assert max(ext) > 32
    _check(ctx, sy, S.SENK_ILU0, 0, 16, colors=True)


@pytest.mark.parametrize("p,bsize", [(1, 8), (2, 4)])
def test_layout_ilup_fill(ctx, p, bsize):
    # ILU(p) fill can couple blocks of one color: the handle falls back to the
    # dependency levels of the factor (block lists), which the layout also takes
    rptr, cind, val = S.gen_convdiff27(10, 10.0)
    sy = abmc_system(rptr, cind, val, bsize)
    _check(ctx, sy, S.SENK_ILUP, p, bsize, colors=True)


@pytest.mark.parametrize("bsize", [1, 4])
def test_layout_natural_order_levels(ctx, bsize):
    rptr, cind, val = S.gen_unstructured(2048, 5)
    N = len(rptr) - 1
    sy = dict(N=N, nc=0, cptr=np.zeros(1, np.int32), val=val, cind=cind, rptr=rptr)
    _check(ctx, sy, S.SENK_ILU0, 0, bsize, colors=False)


@pytest.mark.parametrize("method,n,bsize", [("simple", 32, 8), ("simple", 24, 4), ("connect", 40, 8)])
def test_layout_ticket_order(ctx, method, n, bsize):
    # levels of hundreds of slices: padding every level to whole groups of 64 slices (8
    # streams x 8 waves) adds less than the slices themselves, so the layout numbers its
    # slices in ticket order (32^3 simple: 2 levels of 256 slices, no padding; 24^3
    # simple: 108 -> 128 per level; 40^3 connect: ~100 -> 128).  The 16^3 cases above
    # (a few slices per level) keep slice order and the ticket table.
    rptr, cind, val = S.gen_poisson7(n)
    sy = abmc_system(rptr, cind, val, bsize, method=method)
    _check(ctx, sy, S.SENK_ILU0, 0, bsize, colors=True)


def test_layout_partial_last_slice(ctx):
    # colors whose row counts are not multiples of 64 leave partial slices
    rptr, cind, val = S.gen_poisson7(10, 9, 7)
    sy = abmc_system(rptr, cind, val, 2)
    assert any((np.diff(sy["cptr"]) * 2) % 64)
    _check(ctx, sy, S.SENK_ILU0, 0, 2, colors=True)


@pytest.mark.parametrize("bsize", [4, 8])
def test_sweep_modes_vardiff64(ctx, bsize):
    # 262,144 rows, thousands of slices per color: many waves wait at once.  Every mode
    # gives the per-level result's bits over 12 applications in a row, and the
    # preconditioned solve takes the same iterations in every mode.
    rptr, cind, val = S.gen_vardiff7(64, 7, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, bsize)
    N = sy["N"]
    A, (Ml, Mc) = _handles(ctx, sy, S.SENK_ILU0, 0, bsize, True)
    try:
        rng = np.random.default_rng(11)
        xs = [rng.uniform(-1, 1, N) for _ in range(3)]
        dxs = [(ctx.array(x), ctx.array(x.astype(np.float32))) for x in xs]
        # one output vector per precision for every application, inputs changing in
        # between: a value left over from an earlier application would show
        dy, dy32 = ctx.zeros(N), ctx.array(np.zeros(N, np.float32))
        ref = {}
        for mode in MODES:
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            for rep in range(4):
                for i, x in enumerate(xs):
                    dx, dx32 = dxs[i]
                    S.call("senk_prec_apply_f64", Ml, dx.p, dy.p)
                    S.call("senk_prec_apply_f32", Ml, dx32.p, dy32.p)
                    got = (dy.download(), dy32.download())
                    if (i, 0) not in ref:
                        ref[(i, 0)], ref[(i, 1)] = got
                    assert np.array_equal(got[0], ref[(i, 0)]), (mode, rep, i)
                    assert np.array_equal(got[1], ref[(i, 1)]), (mode, rep, i)
            S.call("senk_prec_check", Ml)
        # the CSR kernels agree with the first mode's bits
        dx, dy = ctx.array(xs[0]), ctx.zeros(N)
        S.call("senk_prec_apply_f64", Mc, dx.p, dy.p)
        assert np.array_equal(dy.download(), ref[(0, 0)])
        # in place (input and output one vector), in every mode
        for mode in MODES:
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            dxy = ctx.array(xs[1])
            S.call("senk_prec_apply_f64", Ml, dxy.p, dxy.p)
            assert np.array_equal(dxy.download(), ref[(1, 0)]), mode
            dxy32 = ctx.array(xs[1].astype(np.float32))
            S.call("senk_prec_apply_f32", Ml, dxy32.p, dxy32.p)
            assert np.array_equal(dxy32.download(), ref[(1, 1)]), mode
        S.call("senk_prec_check", Ml)
        b = np.random.default_rng(5).uniform(-1, 1, N)
        iters = []
        for mode in MODES:
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            db, dxx = ctx.array(b), ctx.zeros(N)
            o = S.solve_opts(solver=S.SENK_BICGSTAB, eps=1e-10, max_iter=2000)
            res = S.senk_result()
            S.call("senk_solve", ctx.h, A, Ml, db.p, dxx.p, C.byref(o), C.byref(res), None, 0)
            assert res.converged
            iters.append((res.iters, dxx.download()))
        for it, x in iters[1:]:
            assert it == iters[0][0]
            assert np.array_equal(x, iters[0][1])
    finally:
        for M in (Ml, Mc):
            S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_spmv_slice_order_keeps_bits(ctx):
    # creating a colored preconditioner reorders how the matrix's SELL slices are handed
    # to the workgroups (by region of the domain); every row is still computed alone, so
    # the products keep their bits, in FP64 and FP32
    rptr, cind, val = S.gen_vardiff7(40, 2, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, 8)
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    x = np.random.default_rng(9).uniform(-1, 1, N)
    dx, dx32 = ctx.array(x), ctx.array(x.astype(np.float32))

    def products():
        dy, dy32 = ctx.zeros(N), ctx.array(np.zeros(N, np.float32))
        S.call("senk_mat_spmv_f64", A, dx.p, dy.p)
        S.call("senk_mat_spmv_f32", A, dx32.p, dy32.p)
        return dy.download(), dy32.download()

    before = products()
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, 1 << S.SENK_FACTOR_F64, 0)
    M = C.c_void_p()
    S.call("senk_prec_create_ilu", ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
           sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], 8, C.byref(opts), C.byref(M))
    try:
        after = products()
        assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
        rows = np.repeat(np.arange(N), np.diff(sy["rptr"]))
        want = np.bincount(rows, weights=sy["val"] * x[sy["cind"]], minlength=N)
        assert np.max(np.abs(after[0] - want)) <= 1e-12 * np.max(np.abs(want))
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_launch_mode_argument(ctx):
    """Modes 0..3 and -1 (the default, 3) are taken, others refused; after -1 both
    precisions still give the bits of the per-level launches."""
    rptr, cind, val = S.gen_poisson7(12)
    sy = abmc_system(rptr, cind, val, 8)
    N = sy["N"]
    A, (Ml, Mc) = _handles(ctx, sy, S.SENK_ILU0, 0, 8, True)
    try:
        for bad in (-2, 4):
            assert S.lib().senk_prec_set_trsv_mode(Ml, bad) == 1  # SENK_ERR_ARG
        x = np.random.default_rng(9).uniform(-1, 1, N)
        outs = {}
        for mode in (0, -1):
            S.call("senk_prec_set_trsv_mode", Ml, mode)
            dx, dy = ctx.array(x), ctx.zeros(N)
            S.call("senk_prec_apply_f64", Ml, dx.p, dy.p)
            dx32, dy32 = ctx.array(x.astype(np.float32)), ctx.array(np.zeros(N, np.float32))
            S.call("senk_prec_apply_f32", Ml, dx32.p, dy32.p)
            outs[mode] = (dy.download(), dy32.download())
        assert np.array_equal(outs[0][0], outs[-1][0]) and np.array_equal(outs[0][1], outs[-1][1])
        S.call("senk_prec_check", Ml)
    finally:
        for M in (Ml, Mc):
            S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
