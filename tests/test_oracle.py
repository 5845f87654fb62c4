"""Pins of the CPU oracle (oracle/senk_oracle.c), CPU only.

Three independent sources pin every oracle function:
  1. the reference itself, compiled from its headers into oracle/_ref (skipped when
     that library is absent, e.g. on a checkout without the reference tree);
  2. golden fixtures in tests/golden/ that the compiled reference produced
     (tests/golden/make_golden.py), which travel with the repository;
  3. mathematics that does not depend on the reference's loops: dense brute force,
     residual identities, exact LU on banded patterns, the fill-path theorem for
     ILU(p), validity + first-fit equality for the coloring, closed forms for
     Ggen/Grot/Trsv, and the true residual of the Krylov drivers.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import dense, oracle_ilu0, oracle_split_ldu, relerr

GOLDEN = Path(__file__).resolve().parent / "golden"
o = O.orc()
ref = O.ref()
needs_ref = pytest.mark.skipif(ref is None, reason="oracle/_ref (compiled reference) not built")

P, i32, f64 = O.P, O.i32, O.f64


def mats():
    return {
        "poisson7_6": S.gen_poisson7(6),
        "convdiff27_5": S.gen_convdiff27(5, 10.0),
        "unstructured_500": S.gen_unstructured(500, 9),
        "elasticity_3": S.gen_elasticity3(3, 4),
        "vardiff7_6": S.gen_vardiff7(6, 3, 4.0, 1.5),
    }


MATS = mats()
RNG = np.random.default_rng(2024)


# ------------------------------------------------------------------ vs compiled reference
@needs_ref
@pytest.mark.parametrize("name", list(MATS))
def test_spmv_and_sptrsv_match_reference_bitwise(name):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    x = RNG.uniform(-1, 1, N)
    a, b = np.zeros(N), np.zeros(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(x), P(a), N)
    ref.ref_spmv_csr_d(P(val), P(cind), P(rptr), P(x), P(b), N)
    assert np.array_equal(a, b)
    v32, x32 = val.astype(np.float32), x.astype(np.float32)
    a32, b32 = np.zeros(N, np.float32), np.zeros(N, np.float32)
    o.or_spmv_csr_f(P(v32), P(cind), P(rptr), P(x32), P(a32), N)
    ref.ref_spmv_csr_f(P(v32), P(cind), P(rptr), P(x32), P(b32), N)
    assert np.array_equal(a32, b32)
    f = oracle_ilu0(val, cind, rptr)
    g = f64(val).copy()
    ref.ref_ilu0_d(P(g), P(cind), P(rptr), N)
    assert np.array_equal(f, g)
    f32 = v32.copy()
    g32 = v32.copy()
    assert o.or_ilu0_f(P(f32), P(cind), P(rptr), N) == 0
    ref.ref_ilu0_f(P(g32), P(cind), P(rptr), N)
    assert np.array_equal(f32, g32)
    L, U = oracle_split_ldu(f, cind, rptr)
    yl, yu, zl, zu = (np.zeros(N) for _ in range(4))
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(yl), N)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(yl), P(yu), N)
    ref.ref_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(zl), N)
    ref.ref_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(zl), P(zu), N)
    assert np.array_equal(yl, zl) and np.array_equal(yu, zu)
    # diag form
    d = np.diag(dense(rptr, cind, val, N)).copy()
    a, b = np.zeros(N), np.zeros(N)
    o.or_spmv_csr_diag_d(P(val), P(cind), P(rptr), P(d), P(x), P(a), N)
    ref.ref_spmv_csr_diag_d(P(val), P(cind), P(rptr), P(d), P(x), P(b), N)
    assert np.array_equal(a, b)


@needs_ref
@pytest.mark.parametrize("name,bsize,method", [("poisson7_6", 4, 1), ("poisson7_6", 8, 0), ("convdiff27_5", 5, 1),
                                               ("unstructured_500", 4, 1), ("unstructured_500", 10, 0),
                                               ("elasticity_3", 9, 1), ("vardiff7_6", 12, 1)])
def test_orderings_and_colored_solves_match_reference(name, bsize, method):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    for isSym in (0, 1):
        sc1, lp1, rp1 = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
        sc2, lp2, rp2 = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
        n1 = o.or_abmc_perm(P(cind), P(rptr), P(sc1), P(lp1), P(rp1), N, bsize, isSym, method)
        n2 = ref.ref_abmc_perm(P(cind), P(rptr), P(sc2), P(lp2), P(rp2), N, bsize, isSym, method)
        assert n1 == n2
        assert np.array_equal(sc1[: n1 + 1], sc2[: n2 + 1])
        assert np.array_equal(lp1, lp2) and np.array_equal(rp1, rp2)
        n1 = o.or_amc_perm(P(cind), P(rptr), P(sc1), P(lp1), P(rp1), N, isSym)
        n2 = ref.ref_amc_perm(P(cind), P(rptr), P(sc2), P(lp2), P(rp2), N, isSym)
        assert n1 == n2 and np.array_equal(lp1, lp2) and np.array_equal(sc1[: n1 + 1], sc2[: n2 + 1])
    n1 = o.or_abmc_perm(P(cind), P(rptr), P(sc1), P(lp1), P(rp1), N, bsize, 0, method)
    v1, c1, r1 = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    v2, c2, r2 = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    o.or_reordering_d(P(v1), P(c1), P(r1), P(lp1), P(rp1), N)
    ref.ref_reordering_d(P(v2), P(c2), P(r2), P(lp1), P(rp1), N)
    assert np.array_equal(v1, v2) and np.array_equal(c1, c2) and np.array_equal(r1, r2)
    # column-only relabel form
    v3, c3 = f64(val).copy(), i32(cind).copy()
    v4, c4 = f64(val).copy(), i32(cind).copy()
    o.or_reordering_rp_d(P(v3), P(c3), P(rptr), P(rp1), N)
    ref.ref_reordering_rp_d(P(v4), P(c4), P(rptr), P(rp1), N)
    assert np.array_equal(v3, v4) and np.array_equal(c3, c4)
    # colored SpTRSV on the reordered ILU(0) factors
    f = oracle_ilu0(v1, c1, r1)
    L, U = oracle_split_ldu(f, c1, r1)
    x = RNG.uniform(-1, 1, N)
    ya, yb = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_abmc_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(ya), N, P(sc1), n1, bsize)
    ref.ref_sptrsv_l_abmc_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(yb), N, P(sc1), n1, bsize)
    assert np.array_equal(ya, yb)
    ua, ub = np.zeros(N), np.zeros(N)
    o.or_sptrsv_u_abmc_d(P(U[0]), P(U[1]), P(U[2]), P(ya), P(ua), N, P(sc1), n1, bsize)
    ref.ref_sptrsv_u_abmc_d(P(U[0]), P(U[1]), P(U[2]), P(yb), P(ub), N, P(sc1), n1, bsize)
    assert np.array_equal(ua, ub)
    # AMC colored on an AMC-reordered copy
    na = o.or_amc_perm(P(cind), P(rptr), P(sc2), P(lp2), P(rp2), N, 0)
    v5, c5, r5 = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    o.or_reordering_d(P(v5), P(c5), P(r5), P(lp2), P(rp2), N)
    L5, U5 = oracle_split_ldu(oracle_ilu0(v5, c5, r5), c5, r5)
    ya, yb = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_amc_d(P(L5[0]), P(L5[1]), P(L5[2]), P(x), P(ya), N, P(sc2), na)
    ref.ref_sptrsv_l_amc_d(P(L5[0]), P(L5[1]), P(L5[2]), P(x), P(yb), N, P(sc2), na)
    assert np.array_equal(ya, yb)
    o.or_sptrsv_u_amc_d(P(U5[0]), P(U5[1]), P(U5[2]), P(ya), P(ua), N, P(sc2), na)
    ref.ref_sptrsv_u_amc_d(P(U5[0]), P(U5[1]), P(U5[2]), P(yb), P(ub), N, P(sc2), na)
    assert np.array_equal(ua, ub)


@needs_ref
@pytest.mark.parametrize("name,bnl,bnw", [("elasticity_3", 3, 3), ("poisson7_6", 8, 1), ("poisson7_6", 2, 2),
                                          ("convdiff27_5", 5, 1), ("unstructured_500", 4, 2)])
def test_block_formats_and_ilub_match_reference(name, bnl, bnw):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    if N % bnl or N % bnw:
        pytest.skip("shape does not divide N")
    br1 = np.zeros(N // bnl + 1, np.int32)
    nb = o.or_csr2bcsr_count(P(cind), P(rptr), P(br1), N, bnl, bnw)
    bc1, bv1 = np.zeros(nb, np.int32), np.zeros(nb * bnl * bnw)
    o.or_csr2bcsr_fill_d(P(val), P(cind), P(rptr), P(br1), P(bv1), P(bc1), N, bnl, bnw)
    br2, bc2, bv2 = np.zeros(N // bnl + 1, np.int32), np.zeros(nb, np.int32), np.zeros(nb * bnl * bnw)
    assert ref.ref_csr2bcsr_d(P(val), P(cind), P(rptr), P(bv2), P(bc2), P(br2), N, bnl, bnw, nb) == nb
    assert np.array_equal(br1, br2) and np.array_equal(bc1, bc2) and np.array_equal(bv1, bv2)
    x = RNG.uniform(-1, 1, N)
    a, b = np.zeros(N), np.zeros(N)
    o.or_spmv_bcsr_d(P(bv1), P(bc1), P(br1), P(x), P(a), N, bnl, bnw)
    if ref.ref_spmv_bcsr_d(P(bv1), P(bc1), P(br1), P(x), P(b), N, bnl, bnw) == 0:
        assert np.array_equal(a, b)
    cap = nb * bnl * bnw
    cv1, cc1, cr1 = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    cv2, cc2, cr2 = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n1 = o.or_bcsr2csr_d(P(bv1), P(bc1), P(br1), P(cv1), P(cc1), P(cr1), N, bnl, bnw)
    n2 = ref.ref_bcsr2csr_d(P(bv1), P(bc1), P(br1), P(cv2), P(cc2), P(cr2), N, bnl, bnw, cap)
    assert n1 == n2 and np.array_equal(cr1, cr2) and np.array_equal(cc1[:n1], cc2[:n1]) and np.array_equal(cv1[:n1], cv2[:n1])
    av, ac, ar = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    bv, bc, brr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n1 = o.or_alloc_block_zero_d(P(val), P(cind), P(rptr), P(av), P(ac), P(ar), N, bnl, bnw, cap)
    n2 = ref.ref_alloc_block_zero_d(P(val), P(cind), P(rptr), P(bv), P(bc), P(brr), N, bnl, bnw, cap)
    assert n1 == n2 and np.array_equal(ar, brr) and np.array_equal(ac[:n1], bc[:n1]) and np.array_equal(av[:n1], bv[:n1])
    # ILUB factors as BCSR and the BCSR triangular solves (only meaningful when the
    # diagonal square is covered by whole blocks: bnl % bnw == 0)
    if bnl % bnw:
        return
    f = oracle_ilu0(av[:n1], ac[:n1], ar)
    L, U = oracle_split_ldu(f, ac[:n1], ar)
    bl = S.csr2bcsr(*L, N, bnl, bnw)
    bu = S.csr2bcsr(*U, N, bnl, bnw)
    ya, yb, za, zb = (np.zeros(N) for _ in range(4))
    o.or_sptrsv_bcsr_l_d(P(bl[0]), P(bl[1]), P(bl[2]), P(x), P(ya), N, bnl, bnw)
    o.or_sptrsv_bcsr_u_d(P(bu[0]), P(bu[1]), P(bu[2]), P(ya), P(za), N, bnl, bnw)
    if ref.ref_sptrsv_bcsr_l_d(P(bl[0]), P(bl[1]), P(bl[2]), P(x), P(yb), N, bnl, bnw) == 0:
        ref.ref_sptrsv_bcsr_u_d(P(bu[0]), P(bu[1]), P(bu[2]), P(yb), P(zb), N, bnl, bnw)
        assert np.array_equal(ya, yb) and np.array_equal(za, zb)
    # BCSR solve == CSR solve on the block-closed pattern (different loop, same math)
    sl, su = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(sl), N)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(sl), P(su), N)
    assert relerr(za, su) <= 1e-13


@needs_ref
@pytest.mark.parametrize("name", list(MATS))
@pytest.mark.parametrize("p", [1, 2, 3])
def test_ilup_symbolic_and_numeric_match_reference(name, p):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    cap = len(val) * 40
    a = [np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)]
    b = [np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)]
    n1 = o.or_alloc_level_zero_d(P(val), P(cind), P(rptr), P(a[0]), P(a[1]), P(a[2]), N, p, cap)
    n2 = ref.ref_alloc_level_zero_d(P(val), P(cind), P(rptr), P(b[0]), P(b[1]), P(b[2]), N, p, cap)
    assert n1 == n2 > 0
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[1][:n1], b[1][:n1]) and np.array_equal(a[0][:n1], b[0][:n1])
    c = [np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)]
    n3 = o.or_ilup_d(P(val), P(cind), P(rptr), P(c[0]), P(c[1]), P(c[2]), N, p, cap)
    n4 = ref.ref_ilup_d(P(val), P(cind), P(rptr), P(b[0]), P(b[1]), P(b[2]), N, p, cap)
    assert n3 == n4
    assert np.array_equal(c[1][:n3], b[1][:n3])
    assert np.array_equal(c[0][:n3], b[0][:n3])  # == treats -0.0 == 0.0 (sign of zero fills)


@needs_ref
@pytest.mark.parametrize("name", ["poisson7_6", "unstructured_500"])
def test_misc_setup_matches_reference(name):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    nnz = len(val)
    assert o.or_check_structure(P(cind), P(rptr), N) == ref.ref_check_structure_d(P(val), P(cind), P(rptr), N) == 1
    for key in range(4):
        outs1 = [np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(nnz),
                 np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(N)]
        outs2 = [np.zeros_like(t) for t in outs1]
        for inv in (0, 1):
            o.or_split_d(P(val), P(cind), P(rptr), *(P(t) for t in outs1), N, key, inv)
            ref.ref_split_d(P(val), P(cind), P(rptr), *(P(t) for t in outs2), N, key, inv)
            for t1, t2 in zip(outs1, outs2):
                if key == 3 and t1 is outs1[3]:
                    continue
                assert np.array_equal(t1, t2), key
    d1, d2 = np.zeros(N), np.zeros(N)
    o.or_get_diag_d(P(val), P(cind), P(rptr), P(d1), N)
    ref.ref_get_diag_d(P(val), P(cind), P(rptr), P(d2), N)
    assert np.array_equal(d1, d2)
    ld = 1 / np.sqrt(np.abs(d1))
    v1, v2 = f64(val).copy(), f64(val).copy()
    o.or_scaling_d(P(v1), P(cind), P(rptr), P(ld), P(ld), N)
    ref.ref_scaling_d(P(v2), P(cind), P(rptr), P(ld), P(ld), N)
    assert np.array_equal(v1, v2)
    cv1, cr1, cc1 = np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32)
    cv2, cr2, cc2 = np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32)
    o.or_csr2csc_d(P(val), P(cind), P(rptr), P(cv1), P(cr1), P(cc1), N, N)
    ref.ref_csr2csc_d(P(val), P(cind), P(rptr), P(cv2), P(cr2), P(cc2), N, N)
    assert np.array_equal(cv1, cv2) and np.array_equal(cr1, cr2) and np.array_equal(cc1, cc2)
    for size in (64, 32, 7):
        ns = (N + size - 1) // size
        sl1, sl2 = np.zeros(ns + 1, np.int32), np.zeros(ns + 1, np.int32)
        tot = o.or_csr2sell_count(P(rptr), P(sl1), size, N)
        sv1, sc1 = np.zeros(tot), np.zeros(tot, np.int32)
        o.or_csr2sell_fill_d(P(val), P(cind), P(rptr), P(sl1), P(sv1), P(sc1), size, N)
        sv2, sc2 = np.zeros(tot), np.zeros(tot, np.int32)
        assert ref.ref_csr2sell_d(P(val), P(cind), P(rptr), P(sv2), P(sc2), P(sl2), size, N, tot) == tot
        assert np.array_equal(sl1, sl2) and np.array_equal(sv1, sv2) and np.array_equal(sc1, sc2)
        if N % size == 0:  # the reference's SpmvSell is only right for full slices
            x = RNG.uniform(-1, 1, N)
            a, b = np.zeros(N), np.zeros(N)
            o.or_spmv_sell_d(P(sv1), P(sc1), P(sl1), size, P(x), P(a), N)
            ref.ref_spmv_sell_d(P(sv1), P(sc1), P(sl1), size, P(x), P(b), N)
            assert np.array_equal(a, b)
    bnum = 4
    v1, c1, r1 = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    v2, c2, r2 = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    n1 = o.or_remove_off_diagonal_d(P(v1), P(c1), P(r1), N, bnum)
    n2 = ref.ref_remove_off_diagonal_d(P(v2), P(c2), P(r2), N, bnum)
    assert n1 == n2 and np.array_equal(r1, r2) and np.array_equal(c1[:n1], c2[:n1]) and np.array_equal(v1[:n1], v2[:n1])
    L, U = oracle_split_ldu(oracle_ilu0(v1[:n1], c1[:n1], r1), c1[:n1], r1)
    x = RNG.uniform(-1, 1, N)
    ya, yb, za, zb = (np.zeros(N) for _ in range(4))
    o.or_sptrsv_l_bj_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(ya), N, bnum)
    ref.ref_sptrsv_l_bj_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(yb), N, bnum)
    o.or_sptrsv_u_bj_d(P(U[0]), P(U[1]), P(U[2]), P(ya), P(za), N, bnum)
    ref.ref_sptrsv_u_bj_d(P(U[0]), P(U[1]), P(U[2]), P(yb), P(zb), N, bnum)
    assert np.array_equal(ya, yb) and np.array_equal(za, zb)
    # padding
    extra = (128 - N % 128) % 128
    pv1, pc1 = np.concatenate([val, np.zeros(extra)]), np.concatenate([cind, np.zeros(extra, np.int32)])
    pr1 = np.concatenate([rptr, np.zeros(extra, np.int32)])
    pv2, pc2, pr2 = pv1.copy(), pc1.copy(), pr1.copy()
    n1 = o.or_padding_d(P(pv1), P(pc1), P(pr1), 128, N)
    n2 = ref.ref_padding_d(P(pv2), P(pc2), P(pr2), 128, N)
    assert n1 == n2 == N + extra
    assert np.array_equal(pv1, pv2) and np.array_equal(pc1, pc2) and np.array_equal(pr1, pr2)


@needs_ref
def test_blas_match_reference():
    N = 10007
    x, y = RNG.uniform(-1, 1, N), RNG.uniform(-1, 1, N)
    for fo, fr, args in (("or_axpy_d", "ref_axpy_d", (0.3,)), ("or_scal_d", "ref_scal_d", (1.7,))):
        a, b = y.copy(), y.copy()
        if fo == "or_axpy_d":
            getattr(o, fo)(0.3, P(x), P(a), N)
            getattr(ref, fr)(0.3, P(x), P(b), N)
        else:
            getattr(o, fo)(1.7, P(a), N)
            getattr(ref, fr)(1.7, P(b), N)
        assert np.array_equal(a, b)
    a, b = y.copy(), y.copy()
    o.or_axpby_d(0.3, P(x), -2.0, P(a), N)
    ref.ref_axpby_d(0.3, P(x), -2.0, P(b), N)
    assert np.array_equal(a, b)
    z1, z2 = np.zeros(N), np.zeros(N)
    o.or_axpyz_d(0.3, P(x), P(y), P(z1), N)
    ref.ref_axpyz_d(0.3, P(x), P(y), P(z2), N)
    assert np.array_equal(z1, z2)
    a, b = y.copy(), y.copy()
    o.or_hadprod_d(P(x), P(a), N)
    ref.ref_hadprod_d(P(x), P(b), N)
    assert np.array_equal(a, b)
    a, b = y.copy(), y.copy()
    o.or_haddiv_d(P(x), P(a), N)
    ref.ref_haddiv_d(P(x), P(b), N)
    assert np.array_equal(a, b)
    # reductions: the reference's OpenMP order depends on the thread count
    s = float(np.sum(np.abs(x * y)))
    assert abs(o.or_dot_d(P(x), P(y), N) - ref.ref_dot_d(P(x), P(y), N)) <= 1e-14 * s
    assert abs(o.or_nrm2_d(P(x), N) - ref.ref_nrm2_d(P(x), N)) <= 1e-14 * np.linalg.norm(x)
    for aa, bb in ((3.0, 4.0), (-1e-3, 2.5), (0.7, -0.2)):
        c1, s1, c2, s2 = C.c_double(), C.c_double(), C.c_double(), C.c_double()
        r1 = o.or_ggen_d(aa, bb, C.byref(c1), C.byref(s1))
        r2 = ref.ref_ggen_d(aa, bb, C.byref(c2), C.byref(s2))
        assert (r1, c1.value, s1.value) == (r2, c2.value, s2.value)
        a1, b1, a2, b2 = C.c_double(aa), C.c_double(bb), C.c_double(aa), C.c_double(bb)
        o.or_grot_d(c1.value, s1.value, C.byref(a1), C.byref(b1))
        ref.ref_grot_d(c1.value, s1.value, C.byref(a2), C.byref(b2))
        assert (a1.value, b1.value) == (a2.value, b2.value)
    n, m = 31, 23
    U = np.triu(RNG.uniform(0.5, 1.5, (n, n))) + 3 * np.eye(n)
    Uc = np.asfortranarray(U).ravel(order="F")
    rhs = RNG.uniform(-1, 1, n)
    x1, x2 = np.zeros(n), np.zeros(n)
    o.or_trsv_d(P(Uc), P(rhs), P(x1), n, m)
    ref.ref_trsv_d(P(Uc), P(rhs), P(x2), n, m)
    assert np.array_equal(x1[:m], x2[:m])


@needs_ref
@pytest.mark.parametrize("name", list(MATS))
def test_float_kernels_and_copy_match_reference_bitwise(name):
    """The FP32 sequential SpTRSV (senk_sparse.hpp:132-167), the FP32 SELL and BCSR SpMV
    (:103-121, :70-91) and blas1::Copy (senk_blas1.hpp:26-30) against the compiled
    reference's own float instantiations, bit for bit."""
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    x = RNG.uniform(-1, 1, N)
    c1, c2 = np.zeros(N), np.full(N, np.nan)
    o.or_copy_d(P(x), P(c1), N)
    ref.ref_copy_d(P(x), P(c2), N)
    assert np.array_equal(c1, c2) and np.array_equal(c1, x)
    x32 = x.astype(np.float32)
    L, U = oracle_split_ldu(oracle_ilu0(val, cind, rptr), cind, rptr)
    L32 = (L[0].astype(np.float32), L[1], L[2])
    U32 = (U[0].astype(np.float32), U[1], U[2])
    yl, yu, zl, zu = (np.zeros(N, np.float32) for _ in range(4))
    o.or_sptrsv_l_f(P(L32[0]), P(L32[1]), P(L32[2]), P(x32), P(yl), N)
    o.or_sptrsv_u_f(P(U32[0]), P(U32[1]), P(U32[2]), P(yl), P(yu), N)
    ref.ref_sptrsv_l_f(P(L32[0]), P(L32[1]), P(L32[2]), P(x32), P(zl), N)
    ref.ref_sptrsv_u_f(P(U32[0]), P(U32[1]), P(U32[2]), P(zl), P(zu), N)
    assert np.array_equal(yl, zl) and np.array_equal(yu, zu)
    # in place, as the ILU drivers call the backward solve (x == y)
    w = yl.copy()
    o.or_sptrsv_u_f(P(U32[0]), P(U32[1]), P(U32[2]), P(w), P(w), N)
    assert np.array_equal(w, yu)
    v32 = val.astype(np.float32)
    for size in (64, 32, 8, 1):
        if N % size:  # the reference's SpmvSell is only right for full slices
            continue
        ns = N // size
        sl = np.zeros(ns + 1, np.int32)
        tot = o.or_csr2sell_count(P(rptr), P(sl), size, N)
        sv, sc = np.zeros(tot), np.zeros(tot, np.int32)
        o.or_csr2sell_fill_d(P(val), P(cind), P(rptr), P(sl), P(sv), P(sc), size, N)
        sv32 = sv.astype(np.float32)
        a, b = np.zeros(N, np.float32), np.full(N, np.nan, np.float32)
        o.or_spmv_sell_f(P(sv32), P(sc), P(sl), size, P(x32), P(a), N)
        ref.ref_spmv_sell_f(P(sv32), P(sc), P(sl), size, P(x32), P(b), N)
        assert np.array_equal(a, b), size
    for bnl, bnw in ((3, 3), (2, 1), (2, 2), (4, 2), (8, 1)):
        if N % bnl or N % bnw:
            continue
        br = np.zeros(N // bnl + 1, np.int32)
        nb = o.or_csr2bcsr_count(P(cind), P(rptr), P(br), N, bnl, bnw)
        bc, bv = np.zeros(nb, np.int32), np.zeros(nb * bnl * bnw)
        o.or_csr2bcsr_fill_d(P(val), P(cind), P(rptr), P(br), P(bv), P(bc), N, bnl, bnw)
        bv32 = bv.astype(np.float32)
        a, b = np.zeros(N, np.float32), np.full(N, np.nan, np.float32)
        o.or_spmv_bcsr_f(P(bv32), P(bc), P(br), P(x32), P(a), N, bnl, bnw)
        assert ref.ref_spmv_bcsr_f(P(bv32), P(bc), P(br), P(x32), P(b), N, bnl, bnw) == 0
        assert np.array_equal(a, b), (bnl, bnw)


@pytest.mark.parametrize("name", list(MATS))
def test_float_kernels_closed_forms(name):
    """The same FP32 functions without the reference: the SELL and BCSR products equal
    the FP32 CSR product bit for bit (their padding and block fill add exact zeros to the
    same ordered row sums) and the FP64 dense product to FP32 rounding; the FP32 solves
    satisfy L y = x and U z = y to FP32 rounding and agree with the FP64 solves."""
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    A = dense(rptr, cind, val, N)
    x = RNG.uniform(-1, 1, N)
    x32, v32 = x.astype(np.float32), val.astype(np.float32)
    csr = np.zeros(N, np.float32)
    o.or_spmv_csr_f(P(v32), P(cind), P(rptr), P(x32), P(csr), N)
    exact = dense(rptr, cind, v32.astype(np.float64), N) @ x32.astype(np.float64)
    scale = np.abs(A) @ np.abs(x)
    assert np.all(np.abs(csr - exact) <= 64 * np.finfo(np.float32).eps * scale)
    for size in (64, 32, 7, 1):
        ns = (N + size - 1) // size
        sl = np.zeros(ns + 1, np.int32)
        tot = o.or_csr2sell_count(P(rptr), P(sl), size, N)
        sv, sc = np.zeros(tot), np.zeros(tot, np.int32)
        o.or_csr2sell_fill_d(P(val), P(cind), P(rptr), P(sl), P(sv), P(sc), size, N)
        a = np.full(N, np.nan, np.float32)
        o.or_spmv_sell_f(P(sv.astype(np.float32)), P(sc), P(sl), size, P(x32), P(a), N)
        assert np.array_equal(a, csr), size
    for bnl, bnw in ((3, 3), (2, 1), (2, 2), (4, 2), (8, 1)):
        if N % bnl or N % bnw:
            continue
        br = np.zeros(N // bnl + 1, np.int32)
        nb = o.or_csr2bcsr_count(P(cind), P(rptr), P(br), N, bnl, bnw)
        bc, bv = np.zeros(nb, np.int32), np.zeros(nb * bnl * bnw)
        o.or_csr2bcsr_fill_d(P(val), P(cind), P(rptr), P(br), P(bv), P(bc), N, bnl, bnw)
        a = np.full(N, np.nan, np.float32)
        o.or_spmv_bcsr_f(P(bv.astype(np.float32)), P(bc), P(br), P(x32), P(a), N, bnl, bnw)
        assert np.array_equal(a, csr), (bnl, bnw)
    f = oracle_ilu0(val, cind, rptr)
    F = dense(rptr, cind, f, N)
    Lm, Um = np.eye(N) + np.tril(F, -1), np.triu(F)
    L, U = oracle_split_ldu(f, cind, rptr)
    L32 = (L[0].astype(np.float32), L[1], L[2])
    U32 = (U[0].astype(np.float32), U[1], U[2])
    yl, yu = np.zeros(N, np.float32), np.zeros(N, np.float32)
    o.or_sptrsv_l_f(P(L32[0]), P(L32[1]), P(L32[2]), P(x32), P(yl), N)
    o.or_sptrsv_u_f(P(U32[0]), P(U32[1]), P(U32[2]), P(yl), P(yu), N)
    eps32 = np.finfo(np.float32).eps
    assert np.linalg.norm(Lm @ yl - x32) <= 1e3 * eps32 * np.linalg.norm(x32)
    assert np.linalg.norm(Um @ yu - yl) <= 1e3 * eps32 * np.linalg.norm(yl) * np.linalg.cond(Um)
    dl, du = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x32.astype(np.float64)), P(dl), N)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(dl), P(du), N)
    assert relerr(yu, du) <= 1e3 * eps32 * np.linalg.cond(Lm @ Um)
    c = np.full(N, np.nan)
    o.or_copy_d(P(x), P(c), N)
    assert np.array_equal(c, x)


@pytest.mark.parametrize("name", ["poisson7_6", "unstructured_500", "convdiff27_5"])
def test_relative_residual_pinned(capfd, name):
    """or_relative_residual_d (senk_test.hpp:11-22: ||b - A x|| / ||b|| in FP64) against
    the line the compiled reference prints ("# test %e") and against numpy's dense form."""
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    A = dense(rptr, cind, val, N)
    b, x = RNG.uniform(-1, 1, N), RNG.uniform(-1, 1, N) * 1e-3
    got = o.or_relative_residual_d(P(val), P(cind), P(rptr), P(b), P(x), N)
    want = np.linalg.norm(b - A @ x) / np.linalg.norm(b)
    assert abs(got - want) <= 1e-13 * want
    if ref is not None:
        capfd.readouterr()
        ref.ref_relative_residual_error(P(val), P(cind), P(rptr), P(b), P(x), N)
        ref.ref_flush()
        line = [ln for ln in capfd.readouterr().out.splitlines() if ln.startswith("# test")][-1]
        assert line == "# test %e" % got, (line, got)


def _ref_iters(capfd, fn):
    capfd.readouterr()
    fn()
    ref.ref_flush()
    out = capfd.readouterr().out
    for line in out.splitlines()[::-1]:
        if line.startswith("I iter ") or line.startswith("# iter "):
            return int(line.split()[2])
    for line in out.splitlines()[::-1]:  # GCR prints "i*m res" per restart only
        parts = line.split()
        if len(parts) == 2 and parts[0].isdigit():
            return int(parts[0])
    raise AssertionError(out[-500:])


@needs_ref
@pytest.mark.parametrize("solver", ["gmres", "bicgstab", "gcr", "abmc_gmres", "ilub_gmres"])
def test_krylov_drivers_match_reference(capfd, solver):
    rptr, cind, val = S.gen_poisson7(10) if solver != "ilub_gmres" else S.gen_elasticity3(5, 2)
    N = len(rptr) - 1
    b = S.gen_uniform(1, N)
    nrm_b = float(np.linalg.norm(b))
    conv = C.c_int()
    eps = 1e-8
    if solver == "abmc_gmres":
        sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
        nc = o.or_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, 4, 0, 1)
        val, cind, rptr = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
        o.or_reordering_d(P(val), P(cind), P(rptr), P(lp), P(rp), N)
        b = b[lp].copy()
    L, U = oracle_split_ldu(oracle_ilu0(val, cind, rptr), cind, rptr)
    xo, xr = np.zeros(N), np.zeros(N)
    if solver in ("gmres", "abmc_gmres"):
        it_o = o.or_gmres_d(P(val), P(cind), P(rptr), C.byref(O.prec_spec(1, L, U)), P(b), P(xo), nrm_b, 100, 30, N,
                            eps, None, 0, C.byref(conv))
        if solver == "gmres":
            it_r = _ref_iters(capfd, lambda: ref.ref_ilu_gmresm_d(P(val), P(cind), P(rptr), *(P(t) for t in L),
                                                                  *(P(t) for t in U), P(b), P(xr), nrm_b, 100, 30, N, eps))
        else:
            it_r = _ref_iters(capfd, lambda: ref.ref_abmc_ilu_gmresm_d(P(val), P(cind), P(rptr), *(P(t) for t in L),
                                                                       *(P(t) for t in U), P(sc), nc, 4, P(b), P(xr),
                                                                       nrm_b, 100, 30, N, eps))
    elif solver == "bicgstab":
        it_o = o.or_bicgstab_d(P(val), P(cind), P(rptr), C.byref(O.prec_spec(1, L, U)), P(b), P(xo), nrm_b, 1000, N,
                               eps, None, 0, C.byref(conv))
        it_r = _ref_iters(capfd, lambda: ref.ref_ilu_bicgstab_d(P(val), P(cind), P(rptr), *(P(t) for t in L),
                                                                *(P(t) for t in U), P(b), P(xr), nrm_b, 1000, N, eps))
    elif solver == "gcr":
        it_o = o.or_gcr_d(P(val), P(cind), P(rptr), C.byref(O.prec_spec(1, L, U)), P(b), P(xo), nrm_b, 100, 30, N, eps,
                          None, 0, C.byref(conv))
        it_r = _ref_iters(capfd, lambda: ref.ref_ilu_gcrm_d(P(val), P(cind), P(rptr), *(P(t) for t in L),
                                                            *(P(t) for t in U), P(b), P(xr), nrm_b, 100, 30, N, eps))
    else:
        bnl = bnw = 3
        av, ac, ar = S.alloc_block_zero(val, cind, rptr, N, bnl, bnw)
        Lb, Ub = oracle_split_ldu(oracle_ilu0(av, ac, ar), ac, ar)
        bl, bu = S.csr2bcsr(*Lb, N, bnl, bnw), S.csr2bcsr(*Ub, N, bnl, bnw)
        it_o = o.or_gmres_d(P(val), P(cind), P(rptr), C.byref(O.prec_spec(2, bl, bu, bnl, bnw)), P(b), P(xo), nrm_b,
                            100, 30, N, eps, None, 0, C.byref(conv))
        it_r = _ref_iters(capfd, lambda: ref.ref_ilub_gmresm_d(P(val), P(cind), P(rptr), *(P(t) for t in bl),
                                                               *(P(t) for t in bu), P(b), P(xr), nrm_b, 100, 30, N, eps,
                                                               bnl, bnw))
    assert conv.value == 1
    assert abs(it_o - it_r) <= max(1, it_r // 20), (it_o, it_r)
    assert relerr(xo, xr) <= 1e-6
    # independent pin: the true residual
    assert o.or_relative_residual_d(P(val), P(cind), P(rptr), P(b), P(xo), N) <= 10 * eps


# ------------------------------------------------------------------ golden fixtures
def _golden():
    # the small-system fixtures (the 32^3 ordering and the MatrixMarket files have tests of
    # their own)
    return sorted(p for p in GOLDEN.glob("*.npz")
                  if not p.name.startswith("poisson7_32_") and p.name != "matrix_market.npz")


@pytest.mark.skipif(not _golden(), reason="no golden fixtures")
@pytest.mark.parametrize("path", _golden(), ids=lambda p: p.stem)
def test_oracle_against_golden(path):
    g = np.load(path, allow_pickle=False)
    rptr, cind, val = g["rptr"], g["cind"], g["val"]
    N = len(rptr) - 1
    x = g["x"]
    y = np.zeros(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(x), P(y), N)
    assert np.array_equal(y, g["spmv"])
    f = oracle_ilu0(val, cind, rptr)
    assert np.array_equal(f, g["ilu0"])
    L, U = oracle_split_ldu(f, cind, rptr)
    yl, yu = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(yl), N)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(yl), P(yu), N)
    assert np.array_equal(yu, g["ilu_apply"])
    for key in [k for k in g.files if k.startswith("abmc_")]:
        bsize, method = (int(t) for t in key.split("_")[1:3])
        sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
        nc = o.or_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, bsize, 0, method)
        assert np.array_equal(np.concatenate([[nc], sc[: nc + 1], lp]), g[key]), key
    for key in [k for k in g.files if k.startswith("ilup_")]:
        p = int(key.split("_")[1])
        cap = len(val) * 40
        ov, oc, orr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
        n = o.or_alloc_level_zero_d(P(val), P(cind), P(rptr), P(ov), P(oc), P(orr), N, p, cap)
        assert np.array_equal(np.concatenate([orr, oc[:n]]), g[key]), key
    if "gmres_iters" in g.files:
        b = g["b"]
        xo = np.zeros(N)
        conv = C.c_int()
        it = o.or_gmres_d(P(val), P(cind), P(rptr), C.byref(O.prec_spec(1, L, U)), P(b), P(xo),
                          float(np.linalg.norm(b)), 100, 30, N, 1e-8, None, 0, C.byref(conv))
        assert abs(it - int(g["gmres_iters"])) <= 1


def test_oracle_poisson32_against_reference_fixture():
    """7-point Poisson 32^3, ABMC bsize 4 connect, ILU(0)-GMRES(30) to 1e-8: the
    reference's ordering arrays and iteration count, as committed fixtures."""
    g = np.load(GOLDEN / "poisson7_32_abmc4.npz", allow_pickle=False)
    rptr, cind, val = S.gen_poisson7(32)
    N = len(rptr) - 1
    sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc = o.or_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, 4, 0, 1)
    assert nc == int(g["num_color"])
    assert np.array_equal(sc[: nc + 1], g["size_color"]) and np.array_equal(lp, g["LP"])
    v, c, r = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    o.or_reordering_d(P(v), P(c), P(r), P(lp), P(rp), N)
    L, U = oracle_split_ldu(oracle_ilu0(v, c, r), c, r)
    xs = S.gen_uniform(1, N)
    b = np.zeros(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(xs), P(b), N)
    bp = b[lp].copy()
    x = np.zeros(N)
    conv = C.c_int()
    it = o.or_gmres_d(P(v), P(c), P(r), C.byref(O.prec_spec(1, L, U)), P(bp), P(x), float(np.linalg.norm(bp)), 200,
                      30, N, 1e-8, None, 0, C.byref(conv))
    assert conv.value == 1 and abs(it - int(g["gmres_iters"])) <= 1


# ------------------------------------------------------------------ mathematics (no reference)
@pytest.mark.parametrize("name", list(MATS))
def test_spmv_brute_force_and_row_sums(name):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    A = dense(rptr, cind, val, N)
    x = RNG.uniform(-1, 1, N)
    y = np.zeros(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(x), P(y), N)
    assert relerr(y, A @ x) <= 1e-14
    ones = np.ones(N)
    o.or_spmv_csr_d(P(val), P(cind), P(rptr), P(ones), P(y), N)
    assert np.allclose(y, A.sum(1), rtol=0, atol=1e-13)


@pytest.mark.parametrize("name", list(MATS))
def test_ilu0_defining_property_and_trsv_residuals(name):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    f = oracle_ilu0(val, cind, rptr)
    F = dense(rptr, cind, f, N)
    A = dense(rptr, cind, val, N)
    Lm, Um = np.eye(N) + np.tril(F, -1), np.triu(F)
    mask = np.zeros((N, N), bool)
    for i in range(N):
        mask[i, cind[rptr[i]:rptr[i + 1]]] = True
    assert np.max(np.abs((Lm @ Um - A)[mask])) <= 1e-12 * np.max(np.abs(A))
    L, U = oracle_split_ldu(f, cind, rptr)
    x = RNG.uniform(-1, 1, N)
    yl, yu = np.zeros(N), np.zeros(N)
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(x), P(yl), N)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(yl), P(yu), N)
    assert np.linalg.norm(Lm @ yl - x) <= 1e-12 * np.linalg.norm(x)
    assert np.linalg.norm(Um @ yu - yl) <= 1e-12 * np.linalg.norm(yl) * np.linalg.cond(Um)


def test_ilu0_exact_on_tridiagonal():
    n = 50
    rng = np.random.default_rng(5)
    lo, di, up = rng.uniform(-1, 0, n - 1), rng.uniform(3, 4, n), rng.uniform(-1, 0, n - 1)
    A = np.diag(di) + np.diag(lo, -1) + np.diag(up, 1)
    rptr = [0]
    cind, val = [], []
    for i in range(n):
        for j in range(max(0, i - 1), min(n, i + 2)):
            cind.append(j)
            val.append(A[i, j])
        rptr.append(len(cind))
    rptr, cind, val = i32(rptr), i32(cind), f64(val)
    f = oracle_ilu0(val, cind, rptr)
    F = dense(rptr, cind, f, n)
    Lm, Um = np.eye(n) + np.tril(F, -1), np.triu(F)
    # no fill on a tridiagonal: ILU(0) is the exact LU (Thomas)
    assert np.allclose(Lm @ Um, A, atol=1e-14)
    b = rng.uniform(-1, 1, n)
    import scipy.linalg as sl
    L, U = oracle_split_ldu(f, cind, rptr)
    y, z = np.zeros(n), np.zeros(n)
    o.or_sptrsv_l_d(P(L[0]), P(L[1]), P(L[2]), P(b), P(y), n)
    o.or_sptrsv_u_d(P(U[0]), P(U[1]), P(U[2]), P(y), P(z), n)
    assert relerr(z, sl.solve(A, b)) <= 1e-13


def _fill_levels_bruteforce(pattern, N, p):
    """Level-of-fill by the fill-path theorem: level(i,j) = (shortest path i -> j through
    intermediate nodes < min(i,j)) - 1, computed over the growing elimination graph."""
    INF = 10 ** 9
    lev = np.full((N, N), INF, dtype=np.int64)
    for i, j in pattern:
        lev[i, j] = 0
    for k in range(N):  # eliminate k: paths through nodes <= k
        for i in range(k + 1, N):
            if lev[i, k] > p:
                continue
            for j in range(k + 1, N):
                if lev[k, j] < INF:
                    lev[i, j] = min(lev[i, j], lev[i, k] + lev[k, j] + 1)
    return lev


@pytest.mark.parametrize("seed", [0, 1, 2])
@pytest.mark.parametrize("p", [0, 1, 2, 5])
def test_ilup_symbolic_fill_path_theorem(seed, p):
    rng = np.random.default_rng(seed)
    N = 24
    M = (rng.random((N, N)) < 0.12) | np.eye(N, dtype=bool)
    rptr = [0]
    cind, val = [], []
    for i in range(N):
        cols = np.nonzero(M[i])[0]
        cind += cols.tolist()
        val += [1.0] * len(cols)
        rptr.append(len(cind))
    rptr, cind, val = i32(rptr), i32(cind), f64(val)
    cap = N * N
    ov, oc, orr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n = o.or_alloc_level_zero_d(P(val), P(cind), P(rptr), P(ov), P(oc), P(orr), N, p, cap)
    got = {(i, int(c)) for i in range(N) for c in oc[orr[i]:orr[i + 1]]}
    lev = _fill_levels_bruteforce([(i, j) for i in range(N) for j in np.nonzero(M[i])[0]], N, p)
    want = {(i, j) for i in range(N) for j in range(N) if lev[i, j] <= p}
    assert got == want
    assert n == len(want)
    if p == 5:
        # large p on this size: the symbolic LU pattern (every fill path admitted)
        lev_inf = _fill_levels_bruteforce([(i, j) for i in range(N) for j in np.nonzero(M[i])[0]], N, 10 ** 6)
        full = {(i, j) for i in range(N) for j in range(N) if lev_inf[i, j] < 10 ** 9}
        assert got <= full


@pytest.mark.parametrize("name", list(MATS))
def test_coloring_valid_and_first_fit(name):
    rptr, cind, _ = MATS[name]
    N = len(rptr) - 1
    ac, ar = np.zeros(2 * len(cind), np.int32), np.zeros(N + 1, np.int32)
    o.or_adjacency(P(cind), P(rptr), P(ac), P(ar), N, 0)
    color = np.zeros(N, np.int32)
    nc = o.or_coloring(P(ac), P(ar), P(color), N)
    adj = [set(ac[ar[i]:ar[i + 1]].tolist()) - {i} for i in range(N)]
    for i in range(N):
        assert all(color[j] != color[i] for j in adj[i])
    ff = np.zeros(N, np.int32)
    for i in range(N):
        used = {ff[j] for j in adj[i] if j < i}
        c = 1
        while c in used:
            c += 1
        ff[i] = c
    assert np.array_equal(color, ff) and nc == ff.max()


@pytest.mark.parametrize("name", list(MATS))
def test_reordering_is_permutation_similarity(name):
    rptr, cind, val = MATS[name]
    N = len(rptr) - 1
    sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    o.or_amc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, 0)
    assert np.array_equal(np.sort(lp), np.arange(N)) and np.array_equal(rp[lp], np.arange(N))
    v, c, r = f64(val).copy(), i32(cind).copy(), i32(rptr).copy()
    o.or_reordering_d(P(v), P(c), P(r), P(lp), P(rp), N)
    A, B = dense(rptr, cind, val, N), dense(r, c, v, N)
    assert np.array_equal(B, A[np.ix_(lp, lp)])  # B[i][RP[c]] = A[LP[i]][c]
    for i in range(N):
        assert np.all(np.diff(c[r[i]:r[i + 1]]) > 0)


def test_givens_and_trsv_closed_forms():
    for a, b in ((3.0, 4.0), (1.0, 0.0), (-2.0, 5.0)):
        c, s = C.c_double(), C.c_double()
        r = o.or_ggen_d(a, b, C.byref(c), C.byref(s))
        assert np.isclose(r, np.hypot(a, b))
        x, y = C.c_double(a), C.c_double(b)
        o.or_grot_d(c.value, s.value, C.byref(x), C.byref(y))
        assert np.isclose(x.value, r) and abs(y.value) <= 1e-15 * r  # rotation zeroes b
    n, m = 12, 9
    rng = np.random.default_rng(3)
    U = np.triu(rng.uniform(0.5, 1.0, (n, n))) + np.eye(n)
    b = rng.uniform(-1, 1, n)
    x = np.zeros(n)
    o.or_trsv_d(P(U.ravel(order="F").copy()), P(b), P(x), n, m)
    assert relerr(U[:m, :m] @ x[:m], b[:m]) <= 1e-14


def test_planted_mistake_is_caught():
    """A deliberately wrong restatement (a dropped pivot division, planted here on the
    CPU) must fail the defining property the ILU(0) pin uses."""
    rptr, cind, val = MATS["convdiff27_5"]
    N = len(rptr) - 1
    f = f64(val).copy()
    for i in range(1, N):  # IKJ without the division by the pivot
        for k in range(rptr[i], rptr[i + 1]):
            ck = cind[k]
            if ck >= i:
                break
            for j in range(k + 1, rptr[i + 1]):
                hit = np.nonzero(cind[rptr[ck]:rptr[ck + 1]] == cind[j])[0]
                if hit.size:
                    f[j] -= f[k] * f[rptr[ck] + hit[0]]
    F, A = dense(rptr, cind, f, N), dense(rptr, cind, val, N)
    Lm, Um = np.eye(N) + np.tril(F, -1), np.triu(F)
    mask = A != 0
    assert np.max(np.abs((Lm @ Um - A)[mask])) > 1e-3
