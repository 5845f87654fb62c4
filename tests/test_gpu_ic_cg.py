"""IC(0)/IC(p) factors and the CG driver on the GPU (the reference announces both,
README_ja.md:44,47, and has neither; the oracle's or_ic0_d / or_cg_d are pinned in
test_ic_cg_oracle.py).

* Factors: the device IC equals the device ILU(0) in L bit for bit (the oracle's ILU(0),
  pinned against the reference) and has U = D L^T exactly (d_i l_ji in FP64); against the
  oracle's independent incomplete LDL^T it agrees to 1e-12; the application is the
  sequential L / U solves.
* CG: FP64 IC(0)-CG in every triangular-solve launch mode and on the CSR kernels,
  iteration count within 5% (at least +-2) of the oracle's IC(0)-CG, FP64 true residual
  at the tolerance, error against x*; IC(1); the mixed forms with an FP32 inner CG.
* Errors: IC of an unsymmetric matrix, CG with the learned correction.
"""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, oracle_ilu0, relerr

pytestmark = pytest.mark.gpu
o = O.orc()


def _system(n=14, bsize=8, seed=3):
    rptr, cind, val = S.gen_vardiff7(n, seed, 4.0, 1.5)
    sy = abmc_system(rptr, cind, val, bsize)
    N = sy["N"]
    xs = np.random.default_rng(seed).uniform(-1, 1, N)
    b = np.zeros(N)
    o.or_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xs), O.P(b), N)
    sy["b"], sy["xs"] = b, xs
    return sy


def _handles(ctx, sy, kind=S.SENK_IC0, p=0, csr_only=0):
    N = sy["N"]
    A = C.c_void_p()
    S.call("senk_mat_create", ctx.h, N, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]), 1 | 2, C.byref(A))
    opts = S.senk_ilu_opts(kind, p, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32), csr_only)
    M = C.c_void_p()
    st = S.lib().senk_prec_create_ilu(ctx.h, A, O.P(sy["rptr"]), O.P(sy["cind"]), O.P(sy["val"]),
                                      sy["cptr"].ctypes.data_as(C.c_void_p), sy["nc"], sy["bsize"], C.byref(opts),
                                      C.byref(M))
    return A, M, st


def _download(M, N):
    ln, un, nl, nu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    S.call("senk_prec_factor_nnz", M, C.byref(ln), C.byref(un), C.byref(nl), C.byref(nu))
    lr, lc, lv = np.zeros(N + 1, np.int32), np.zeros(ln.value, np.int32), np.zeros(ln.value)
    ur, uc, uv = np.zeros(N + 1, np.int32), np.zeros(un.value, np.int32), np.zeros(un.value)
    S.call("senk_prec_download", M, O.P(lr), O.P(lc), O.P(lv), O.P(ur), O.P(uc), O.P(uv))
    return (lv, lc, lr), (uv, uc, ur)


def _oracle_ic(sy, p):
    N = sy["N"]
    if p:
        fv, fc, fr = S.alloc_level_zero(sy["val"], sy["cind"], sy["rptr"], N, p)
    else:
        fv, fc, fr = sy["val"], sy["cind"], sy["rptr"]
    f = np.ascontiguousarray(fv, dtype=np.float64).copy()
    assert o.or_ic0_d(O.P(f), O.P(O.i32(fc)), O.P(O.i32(fr)), N) == 0
    g = oracle_ilu0(fv, fc, fr)
    rows = np.repeat(np.arange(N), np.diff(fr))
    pivots = np.zeros(N)
    pivots[rows[fc == rows]] = g[fc == rows]  # the ILU's pivots, before inversion
    return S.split(f, fc, fr, N, "L-DU", True), S.split(g, fc, fr, N, "L-DU", True), pivots


@pytest.mark.parametrize("p", [0, 1])
def test_ic_factors(ctx, p):
    sy = _system(10, 8)
    N = sy["N"]
    A, M, st = _handles(ctx, sy, S.SENK_ICP if p else S.SENK_IC0, p)
    assert st == 0
    try:
        (lv, lc, lr), (uv, uc, ur) = _download(M, N)
        ic, ilu, piv = _oracle_ic(sy, p)
        # L: the ILU(0) multipliers, bit for bit
        assert np.array_equal(lr, ilu["L"][2]) and np.array_equal(lc, ilu["L"][1])
        assert np.array_equal(lv, ilu["L"][0])
        assert np.array_equal(ur, ilu["U"][2]) and np.array_equal(uc, ilu["U"][1])
        # U: the inverted pivots first (the ILU's), then d_i l_ji exactly
        assert np.array_equal(uv[ur[:-1]], ilu["U"][0][ur[:-1]])
        lrows = np.repeat(np.arange(N), np.diff(lr))
        Lmap = {(int(i), int(j)): v for i, j, v in zip(lrows, lc, lv)}
        urows = np.repeat(np.arange(N), np.diff(ur))
        off = np.ones(len(uc), bool)
        off[ur[:-1]] = False
        want = np.array([piv[i] * Lmap[(int(j), int(i))] for i, j in zip(urows[off], uc[off])])
        assert np.array_equal(uv[off], want)
        # against the oracle's incomplete LDL^T
        assert relerr(lv, ic["L"][0]) <= 1e-12 and relerr(uv, ic["U"][0]) <= 1e-12
        # the application equals the sequential solves with these factors
        x = np.random.default_rng(1).uniform(-1, 1, N)
        yl, yu = np.zeros(N), np.zeros(N)
        o.or_sptrsv_l_d(O.P(lv), O.P(lc), O.P(lr), O.P(x), O.P(yl), N)
        o.or_sptrsv_u_d(O.P(uv), O.P(uc), O.P(ur), O.P(yl), O.P(yu), N)
        dx, dy = ctx.array(x), ctx.zeros(N)
        S.call("senk_prec_apply_f64", M, dx.p, dy.p)
        assert np.array_equal(dy.download(), yu)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def _solve(ctx, A, M, sy, **kw):
    db, dx = ctx.array(sy["b"]), ctx.zeros(sy["N"])
    o_ = S.solve_opts(**kw)
    res = S.senk_result()
    st = S.lib().senk_solve(ctx.h, A, M, db.p, dx.p, C.byref(o_), C.byref(res), None, 0)
    return st, res, dx.download()


def _oracle_cg(sy, p, eps):
    N = sy["N"]
    ic, _, _ = _oracle_ic(sy, p)
    x = np.zeros(N)
    conv = C.c_int()
    it = o.or_cg_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(1, ic["L"], ic["U"])),
                   O.P(sy["b"]), O.P(x), float(np.linalg.norm(sy["b"])), 5000, N, eps, None, 0, C.byref(conv))
    assert conv.value == 1
    return it


def _close(it, ref):
    return abs(it - ref) <= max(2, int(np.ceil(0.05 * ref)))


@pytest.mark.parametrize("mode", [-1, 0, 1, 2, 3, "csr"])
def test_ic0_cg_fp64(ctx, mode):
    sy = _system(16, 8)
    eps = 1e-10
    it_o = _oracle_cg(sy, 0, eps)
    A, M, st = _handles(ctx, sy, csr_only=1 if mode == "csr" else 0)
    assert st == 0
    try:
        if mode != "csr":
            S.call("senk_prec_set_trsv_mode", M, mode)
        print(f"cg mode {mode}", flush=True)
        st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=eps, max_iter=5000)
        assert st == 0 and res.converged == 1 and res.status == 0
        assert _close(res.iters, it_o), (res.iters, it_o)
        assert res.true_relres <= 1.5 * eps
        assert relerr(x, sy["xs"]) <= 1e-7
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_ic1_cg_and_fewer_iterations(ctx):
    sy = _system(16, 8, seed=5)
    eps = 1e-10
    its = {}
    for p in (0, 1):
        A, M, st = _handles(ctx, sy, S.SENK_ICP if p else S.SENK_IC0, p)
        assert st == 0
        try:
            st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=eps, max_iter=5000)
            assert st == 0 and res.converged == 1 and res.true_relres <= 1.5 * eps
            assert _close(res.iters, _oracle_cg(sy, p, eps))
            its[p] = res.iters
        finally:
            S.lib().senk_prec_destroy(M)
            S.lib().senk_mat_destroy(A)
    assert its[1] < its[0]


@pytest.mark.parametrize("mixed", [1, 2])
def test_mixed_forms_with_inner_cg(ctx, mixed):
    sy = _system(20, 8, seed=2)
    A, M, st = _handles(ctx, sy)
    assert st == 0
    try:
        st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-10, mixed=mixed, inner_eps=1e-3,
                            inner_max=60, outer_m=20, max_iter=2000)
        assert st == 0 and res.converged == 1 and res.status == 0
        assert res.true_relres <= 1.5e-10
        assert relerr(x, sy["xs"]) <= 1e-7
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_errors(ctx):
    rptr, cind, val = S.gen_convdiff27(8, 10.0)  # nonsymmetric values
    sy = abmc_system(rptr, cind, val, 8)
    A, M, st = _handles(ctx, sy)
    assert st == 1  # SENK_ERR_ARG
    assert b"not symmetric" in S.lib().senk_last_error()
    S.lib().senk_mat_destroy(A)
    sy = _system(8, 8)
    sy["b"] = np.ones(sy["N"])
    A, M, st = _handles(ctx, sy)
    assert st == 0
    try:
        st, res, _ = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-8, max_iter=100, use_learned=1)
        assert st == 1  # no learned correction with CG (and none attached)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("kind", [S.SENK_IC0, S.SENK_ILU0])
@pytest.mark.parametrize("mode", [3, 2, 1, 0])
def test_split_inner_cg(ctx, kind, mode):
    """The split form of the FP32 inner CG (CG on L^-1 A L^-T preconditioned by D^-1,
    one U and one L sweep per step, no SpMV) against the plain inner CG: both converge to
    the FP64 tolerance, with iteration counts close to each other; a repeated solve
    repeats its iterations and the plain application is unchanged after split solves."""
    sy = _system(20, 8, seed=4)
    N = sy["N"]
    A, M, st = _handles(ctx, sy, kind)
    assert st == 0
    try:
        S.call("senk_prec_set_trsv_mode", M, mode)
        xr = np.random.default_rng(6).uniform(-1, 1, N).astype(np.float32)
        dxr, dyr = ctx.array(xr), ctx.zeros(N, np.float32)
        S.call("senk_prec_apply_f32", M, dxr.p, dyr.p)
        z0 = dyr.download()
        its = {}
        for mixed in (2, 1):
            for split in (1, 0):
                S.call("senk_prec_set_split", M, split)
                st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-10, mixed=mixed, inner_eps=1e-3,
                                    inner_max=60, outer_m=20, max_iter=3000)
                assert st == 0 and res.converged == 1 and res.status == 0, (mixed, split)
                assert res.true_relres <= 1.5e-10, (mixed, split, res.true_relres)
                assert relerr(x, sy["xs"]) <= 1e-7
                its[(mixed, split)] = res.iters
            assert its[(mixed, 1)] <= 1.3 * its[(mixed, 0)] + 10, its
        S.call("senk_prec_set_split", M, 1)
        st, r2, x2 = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-10, mixed=2, inner_eps=1e-3, inner_max=60,
                            outer_m=20, max_iter=3000)
        assert r2.iters == its[(2, 1)]
        S.call("senk_prec_apply_f32", M, dxr.p, dyr.p)
        assert np.array_equal(dyr.download(), z0)
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("kind", [S.SENK_IC0, S.SENK_ILU0])
@pytest.mark.parametrize("inner", [(60, 1e-3), (7, 1e-3)])
@pytest.mark.parametrize("grid", [(18, 8), (13, 1)])
def test_split_inner_cg_modes_bit_identical(ctx, kind, inner, grid):
    """The split inner CG forms its vector passes inside the sweeps in the single-launch
    modes (w = u - E t after the U solve, q = D t + L^-1 w after the L solve, and the
    previous step's yd, r, u update before the U solve reads u) and as separate passes in
    the per-level mode, with the same arithmetic: every mode gives the same iterations and
    the same solution bit for bit.  Inner solves that stop on their tolerance and inner
    solves cut at the step limit (7 steps) both end with the last step's yd update.  13^3
    rows in blocks of 1 (N = 2197, not a multiple of 4): the passes' last group of four
    takes its missing places as zeros, and divides by 1 there, not by 0."""
    sy = _system(grid[0], grid[1], seed=9)
    A, M, st = _handles(ctx, sy, kind)
    assert st == 0
    try:
        S.call("senk_prec_set_split", M, 1)
        out = {}
        for mode in (0, 1, 2, 3):
            S.call("senk_prec_set_trsv_mode", M, mode)
            print(f"split cg mode {mode} inner {inner}", flush=True)
            st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-10, mixed=2, inner_max=inner[0],
                                inner_eps=inner[1], outer_m=30, max_iter=3000)
            assert st == 0 and res.converged == 1 and res.status == 0, mode
            assert res.true_relres <= 1.5e-10 and relerr(x, sy["xs"]) <= 1e-7
            out[mode] = (res.iters, res.restarts, x)
        for mode in (1, 2, 3):
            assert out[mode][:2] == out[0][:2], (mode, out[mode][:2], out[0][:2])
            assert np.array_equal(out[mode][2], out[0][2]), mode
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("mode", [3, "csr"])
def test_cg_zero_right_hand_side(ctx, mode):
    """b = 0 from x0 = 0: (r, z) = 0 at the start, so CG stops at once as converged and x
    stays exactly zero (the first step would otherwise divide 0 by 0)."""
    sy = _system(8, 8)
    sy["b"] = np.zeros(sy["N"])
    A, M, st = _handles(ctx, sy, csr_only=1 if mode == "csr" else 0)
    assert st == 0
    try:
        st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_CG, eps=1e-10, max_iter=100)
        assert st == 0 and res.status == 0 and res.converged == 1 and res.iters == 0
        assert np.array_equal(x, np.zeros(sy["N"]))
        S.call("senk_prec_check", M)
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


@pytest.mark.parametrize("mixed", [2, 1])
def test_split_inner_bicgstab_with_ic_factors(ctx, mixed):
    """IC(0) factors also carry the split-form vectors (U = D L^T equals A's upper part to
    rounding on a triangle-free graph): the split inner BiCGStab converges with them to the
    FP64 tolerance, as with ILU(0) factors, in a count close to the plain inner solve's."""
    sy = _system(20, 8, seed=7)
    A, M, st = _handles(ctx, sy)
    assert st == 0
    try:
        its = {}
        for split in (1, 0):
            S.call("senk_prec_set_split", M, split)
            st, res, x = _solve(ctx, A, M, sy, solver=S.SENK_BICGSTAB, eps=1e-10, mixed=mixed, inner_eps=1e-3,
                                inner_max=40, outer_m=20, max_iter=3000)
            assert st == 0 and res.converged == 1 and res.status == 0, split
            assert res.true_relres <= 1.5e-10 and relerr(x, sy["xs"]) <= 1e-7
            its[split] = res.iters
        assert its[1] <= 1.3 * its[0] + 10, its
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)


def test_ic0_after_jacobi_scaling(ctx):
    """The reference's workflow scales A by D^-1/2 on both sides before factorizing
    (main.cpp:66-73, Scaling senk_matrix.hpp:406-416: ld_i * a_ij * rd_j); the products'
    rounding leaves a_ij and a_ji unequal in the last bit for many entries.  IC(0) accepts
    that matrix (symmetric to a few ulps), factors the symmetric matrix of its lower
    triangle (L bit for bit the oracle's ILU(0) of that matrix, U = D L^T), and FP64 CG and
    the split inner CG converge on the scaled system."""
    sy = _system(16, 8, seed=6)
    N = sy["N"]
    d = np.zeros(N)
    o.or_get_diag_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(d), N)
    s = 1 / np.sqrt(np.abs(d))
    v = sy["val"].copy()
    o.or_scaling_d(O.P(v), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(s), O.P(s), N)
    rows = np.repeat(np.arange(N), np.diff(sy["rptr"]))
    T = {(int(c), int(r)): x for r, c, x in zip(rows, sy["cind"], v)}
    asym = sum(T[(int(r), int(c))] != x for r, c, x in zip(rows, sy["cind"], v))
    assert asym > 0  # the scaled matrix is not bit-symmetric
    sc = dict(sy, val=v)
    xs = sy["xs"] / s  # the scaled system's solution for the same b' = D^-1/2 b
    sc["b"] = sy["b"] * s
    sc["xs"] = xs
    A, M, st = _handles(ctx, sc)
    assert st == 0, S.lib().senk_last_error()
    try:
        (lv, lc, lr), (uv, uc, ur) = _download(M, N)
        mirrored = v.copy()
        for e, (r, c) in enumerate(zip(rows, sy["cind"])):
            if c > r:
                mirrored[e] = T[(int(r), int(c))]
        ilu = S.split(oracle_ilu0(mirrored, sy["cind"], sy["rptr"]), sy["cind"], sy["rptr"], N, "L-DU", True)
        assert np.array_equal(lv, ilu["L"][0]) and np.array_equal(uv[ur[:-1]], ilu["U"][0][ur[:-1]])
        for kw in (dict(solver=S.SENK_CG, eps=1e-10, max_iter=5000),
                   dict(solver=S.SENK_CG, eps=1e-10, mixed=2, inner_eps=1e-3, inner_max=60, outer_m=20,
                        max_iter=3000)):
            st, res, x = _solve(ctx, A, M, sc, **kw)
            assert st == 0 and res.converged == 1 and res.status == 0, kw
            assert res.true_relres <= 1.5e-10
            assert relerr(x, xs) <= 1e-7
    finally:
        S.lib().senk_prec_destroy(M)
        S.lib().senk_mat_destroy(A)
