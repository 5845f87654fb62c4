"""The library's host-side integer setup against the oracle and the reference's
fixtures, CPU only (no device calls).  Integer outputs must be bit-identical."""
import ctypes as C
import json
from pathlib import Path

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import dense

GOLDEN = Path(__file__).resolve().parent / "golden"
o = O.orc()
P = O.P

CASES = {
    "poisson7_7": lambda: S.gen_poisson7(7),
    "convdiff27_6": lambda: S.gen_convdiff27(6, 10.0),
    "unstructured_1200": lambda: S.gen_unstructured(1200, 2),
    "elasticity_4": lambda: S.gen_elasticity3(4, 1),
    "vardiff7_8": lambda: S.gen_vardiff7(8, 5, 4.0, 1.5),
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("bsize,method", [(1, "connect"), (4, "connect"), (8, "simple"), (16, "connect")])
@pytest.mark.parametrize("isSym", [False, True])
def test_abmc_bit_exact(name, bsize, method, isSym):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    if N % bsize:
        pytest.skip("N % bsize")
    nc, sc, LP, RP = S.abmc_permutation(cind, rptr, N, bsize, isSym, method)
    sc2, lp2, rp2 = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc2 = o.or_abmc_perm(P(cind), P(rptr), P(sc2), P(lp2), P(rp2), N, bsize, int(isSym), int(method == "connect"))
    assert nc == nc2
    assert np.array_equal(sc, sc2[: nc + 1]) and np.array_equal(LP, lp2) and np.array_equal(RP, rp2)


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("isSym", [False, True])
def test_amc_bit_exact(name, isSym):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    nc, sc, LP, RP = S.amc_permutation(cind, rptr, N, isSym)
    sc2, lp2, rp2 = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc2 = o.or_amc_perm(P(cind), P(rptr), P(sc2), P(lp2), P(rp2), N, int(isSym))
    assert nc == nc2 and np.array_equal(sc, sc2[: nc + 1]) and np.array_equal(LP, lp2)


def test_abmc_divisibility_error():
    rptr, cind, val = S.gen_poisson7(5)
    with pytest.raises(S.SenkError) as e:
        S.abmc_permutation(cind, rptr, len(rptr) - 1, 4)
    assert e.value.status == 4


@pytest.mark.parametrize("name", list(CASES))
def test_reordering_split_scaling_bit_exact(name):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    nc, sc, LP, RP = S.amc_permutation(cind, rptr, N, False)
    v, c, r = S.reordering(val, cind, rptr, LP, RP, N)
    v2, c2, r2 = val.copy(), cind.copy(), rptr.copy()
    o.or_reordering_d(P(v2), P(c2), P(r2), P(LP), P(RP), N)
    assert np.array_equal(v, v2) and np.array_equal(c, c2) and np.array_equal(r, r2)
    v3, c3 = S.reordering_rp(val, cind, rptr, RP, N)
    v4, c4 = val.copy(), cind.copy()
    o.or_reordering_rp_d(P(v4), P(c4), P(rptr), P(RP), N)
    assert np.array_equal(v3, v4) and np.array_equal(c3, c4)
    nnz = len(val)
    for key, k in (("LD-U", 0), ("L-DU", 1), ("L-D-U", 2), ("LU-D", 3)):
        for inv in (False, True):
            got = S.split(val, cind, rptr, N, key, inv)
            outs = [np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(nnz),
                    np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(N)]
            o.or_split_d(P(val), P(cind), P(rptr), *(P(t) for t in outs), N, k, int(inv))
            lv, lc, lr = got["L"]
            assert np.array_equal(lr, outs[2]) and np.array_equal(lc, outs[1][: lr[-1]]) and np.array_equal(lv, outs[0][: lr[-1]])
            if "U" in got:
                uv, uc, ur = got["U"]
                assert np.array_equal(ur, outs[5]) and np.array_equal(uc, outs[4][: ur[-1]]) and np.array_equal(uv, outs[3][: ur[-1]])
            if "diag" in got:
                assert np.array_equal(got["diag"], outs[6])
    with pytest.raises(S.SenkError):
        S.split(val, cind, rptr, N, "BAD", True)


@pytest.mark.parametrize("name", list(CASES))
def test_get_diag_scaling_remove_off_diagonal_bit_exact(name):
    """senk_get_diag / senk_scaling / senk_remove_off_diagonal (senk_matrix.hpp:391-416,
    :668-689) against or_get_diag_d / or_scaling_d / or_remove_off_diagonal_d, bit for
    bit, with main.cpp:66-73's scale D^-1/2 and with distinct left and right scales."""
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    pd = C.c_void_p()
    S.call("senk_get_diag", S._ptr(val), S._ptr(cind), S._ptr(rptr), C.byref(pd), N)
    d = S._take(pd, N, np.float64)
    d2 = np.full(N, np.nan)
    o.or_get_diag_d(P(val), P(cind), P(rptr), P(d2), N)
    assert np.array_equal(d, d2)
    assert np.array_equal(d, np.diag(dense(rptr, cind, val, N)))
    rng = np.random.default_rng(11)
    for ld, rd in ((1 / np.sqrt(np.abs(d)), 1 / np.sqrt(np.abs(d))), (rng.uniform(0.5, 2, N), rng.uniform(0.5, 2, N))):
        v1, v2 = val.copy(), val.copy()
        S.call("senk_scaling", S._ptr(v1), S._ptr(cind), S._ptr(rptr), S._ptr(ld), S._ptr(rd), N)
        o.or_scaling_d(P(v2), P(cind), P(rptr), P(ld), P(rd), N)
        assert np.array_equal(v1, v2)
    for bnum in (1, 2, 4, 7):  # 7 leaves a short last block on most of these sizes
        v, c, r = S.remove_off_diagonal(val, cind, rptr, N, bnum)
        v2, c2, r2 = val.copy(), cind.copy(), rptr.copy()
        n = o.or_remove_off_diagonal_d(P(v2), P(c2), P(r2), N, bnum)
        assert len(v) == n and np.array_equal(r, r2) and np.array_equal(c, c2[:n]) and np.array_equal(v, v2[:n])
        # closed form: the entries inside the bnum diagonal blocks of N/bnum rows
        rows = np.repeat(np.arange(N), np.diff(rptr))
        keep = rows // (N // bnum) == cind // (N // bnum)
        assert np.array_equal(c, cind[keep]) and np.array_equal(v, val[keep])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("p", [1, 2])
def test_alloc_level_zero_bit_exact(name, p):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    v, c, r = S.alloc_level_zero(val, cind, rptr, N, p)
    cap = len(val) * 40
    ov, oc, orr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n = o.or_alloc_level_zero_d(P(val), P(cind), P(rptr), P(ov), P(oc), P(orr), N, p, cap)
    assert np.array_equal(r, orr) and np.array_equal(c, oc[:n]) and np.array_equal(v, ov[:n])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("bnl,bnw", [(2, 1), (2, 2), (3, 3), (4, 2), (8, 1)])
def test_block_conversions_bit_exact(name, bnl, bnw):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    if N % bnl or N % bnw:
        pytest.skip("shape")
    bv, bc, br = S.csr2bcsr(val, cind, rptr, N, bnl, bnw)
    br2 = np.zeros(N // bnl + 1, np.int32)
    nb = o.or_csr2bcsr_count(P(cind), P(rptr), P(br2), N, bnl, bnw)
    bc2, bv2 = np.zeros(nb, np.int32), np.zeros(nb * bnl * bnw)
    o.or_csr2bcsr_fill_d(P(val), P(cind), P(rptr), P(br2), P(bv2), P(bc2), N, bnl, bnw)
    assert np.array_equal(br, br2) and np.array_equal(bc, bc2) and np.array_equal(bv, bv2)
    cv, cc, cr = S.bcsr2csr(bv, bc, br, N, bnl, bnw)
    cap = nb * bnl * bnw
    cv2, cc2, cr2 = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n = o.or_bcsr2csr_d(P(bv), P(bc), P(br), P(cv2), P(cc2), P(cr2), N, bnl, bnw)
    assert np.array_equal(cr, cr2) and np.array_equal(cc, cc2[:n]) and np.array_equal(cv, cv2[:n])
    av, ac, ar = S.alloc_block_zero(val, cind, rptr, N, bnl, bnw)
    av2, ac2, ar2 = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    n = o.or_alloc_block_zero_d(P(val), P(cind), P(rptr), P(av2), P(ac2), P(ar2), N, bnl, bnw, cap)
    assert np.array_equal(ar, ar2) and np.array_equal(ac, ac2[:n]) and np.array_equal(av, av2[:n])


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("size", [64, 32, 5])
def test_sell_packing_bit_exact(name, size):
    rptr, cind, val = CASES[name]()
    N = len(rptr) - 1
    sv, sc, sl = S.csr2sell(val, cind, rptr, size, N)
    ns = (N + size - 1) // size
    sl2 = np.zeros(ns + 1, np.int32)
    tot = o.or_csr2sell_count(P(rptr), P(sl2), size, N)
    sv2, sc2 = np.zeros(tot), np.zeros(tot, np.int32)
    o.or_csr2sell_fill_d(P(val), P(cind), P(rptr), P(sl2), P(sv2), P(sc2), size, N)
    assert np.array_equal(sl, sl2) and np.array_equal(sc, sc2) and np.array_equal(sv, sv2)
    # SELL-C-sigma with sigma = 1 and full slices reproduces Csr2Sell's arrays
    h = S.csr2sellcs(val, cind, rptr, N, size, 1)
    try:
        if N % size == 0:
            n = h.nnz_padded
            sp = np.ctypeslib.as_array((C.c_int * (h.nslices + 1)).from_address(h.slice_ptr))
            ci = np.ctypeslib.as_array((C.c_int * n).from_address(h.cind))
            va = np.ctypeslib.as_array((C.c_double * n).from_address(h.val64))
            pe = np.ctypeslib.as_array((C.c_int * (h.nslices * size)).from_address(h.perm))
            assert np.array_equal(sp, sl) and np.array_equal(ci, sc) and np.array_equal(va, sv)
            assert np.array_equal(pe, np.arange(N))
    finally:
        S.lib().senk_sellcs_free_host(C.byref(h))


def test_sellcs_sigma_sort_is_stable_descending():
    rptr, cind, val = S.gen_unstructured(1000, 8)
    N = len(rptr) - 1
    h = S.csr2sellcs(val, cind, rptr, N, 64, 256)
    try:
        pe = np.ctypeslib.as_array((C.c_int * (h.nslices * 64)).from_address(h.perm)).copy()
        lens = np.diff(rptr)
        for w0 in range(0, N, 256):
            idx = pe[w0:min(N, w0 + 256)]
            assert sorted(idx.tolist()) == list(range(w0, min(N, w0 + 256)))
            ln = lens[idx]
            assert np.all(np.diff(ln) <= 0)
            for L in np.unique(ln):  # stable: equal lengths keep row order
                sel = idx[ln == L]
                assert np.all(np.diff(sel) > 0)
        assert np.all(pe[N:] == -1)
    finally:
        S.lib().senk_sellcs_free_host(C.byref(h))


def test_padding_and_check_structure():
    rptr, cind, val = S.gen_poisson7(5)
    N = len(rptr) - 1
    v, c, r, n = S.padding(val, cind, rptr, 128, N)
    assert n == 128 and len(r) == 129
    assert np.array_equal(v[len(val):], np.ones(n - N)) and np.array_equal(c[len(val):], np.arange(N, n))
    assert S.check_structure(c, r, n)
    bad = cind.copy()
    bad[rptr[3]:rptr[4]] = np.where(bad[rptr[3]:rptr[4]] == 3, 2, bad[rptr[3]:rptr[4]])
    assert not S.check_structure(bad, rptr, N)


def test_csr2csc():
    rptr, cind, val = S.gen_unstructured(700, 3)
    N = len(rptr) - 1
    cv, cr, cc = S.csr2csc(val, cind, rptr, N, N)
    A = dense(rptr, cind, val, N)
    At = dense(cc, cr, cv, N)
    assert np.array_equal(At, A.T)


def test_poisson32_ordering_matches_reference_fixture():
    g = np.load(GOLDEN / "poisson7_32_abmc4.npz", allow_pickle=False)
    rptr, cind, val = S.gen_poisson7(32)
    N = len(rptr) - 1
    nc, sc, LP, RP = S.abmc_permutation(cind, rptr, N, 4, False, "connect")
    assert nc == int(g["num_color"]) and np.array_equal(sc, g["size_color"]) and np.array_equal(LP, g["LP"])


def _sha(*arrays):
    import hashlib

    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


@pytest.mark.parametrize("name,gen,bsize", [
    ("convdiff27_128_abmc16", lambda: S.gen_convdiff27(128, 10.0), 16),
    ("vardiff7_128_abmc16", lambda: S.gen_vardiff7(128, 1, 8.0, 1.5), 16),
    ("elasticity3_96_abmc48", lambda: S.gen_elasticity3(96, 1), 48),
    ("unstructured_4194304_abmc16", lambda: S.gen_unstructured(4194304, 1), 16),
])
def test_full_size_ordering_digest(name, gen, bsize):
    """The benchmark systems' orderings (and the elasticity system's ILU(1) pattern) against the
    reference's digests at full size."""
    path = GOLDEN / f"{name}.json"
    if not path.exists():
        pytest.skip("digest fixture not generated")
    meta = json.loads(path.read_text())
    rptr, cind, val = gen()
    N = len(rptr) - 1
    assert _sha(rptr, cind, val) == meta["sha256_matrix"]
    nc, sc, LP, RP = S.abmc_permutation(cind, rptr, N, bsize, bool(meta["isSym"]), "connect")
    assert nc == meta["num_color"]
    assert _sha(LP) == meta["sha256_LP"] and _sha(RP) == meta["sha256_RP"] and _sha(sc) == meta["sha256_size_color"]
    if "sha256_ilu1_pattern" in meta:
        v, c, r = S.reordering(val, cind, rptr, LP, RP, N)
        del val, cind
        fv, fc, fr = S.alloc_level_zero(v, c, r, N, 1)
        assert len(fc) == meta["ilu1_nnz"]
        assert _sha(fr, fc) == meta["sha256_ilu1_pattern"]


def test_generators_closed_forms():
    n = 6
    rptr, cind, val = S.gen_poisson7(n)
    N = n ** 3
    assert len(val) == 7 * n ** 3 - 6 * n ** 2
    A = dense(rptr, cind, val, N)
    assert np.array_equal(A, A.T)
    deg = (A != 0).sum(1) - 1
    assert np.allclose(A.sum(1), 6 - deg)
    rptr, cind, val = S.gen_convdiff27(n, 10.0)
    assert len(val) == (3 * n - 2) ** 3
    A = dense(rptr, cind, val, N)
    assert np.array_equal(A != 0, (A != 0).T)  # symmetric pattern
    h = 1.0 / (n + 1)
    S_ = (A + A.T) / 2
    K = (A - A.T) / 2
    assert np.allclose(np.diag(A), 26 / 9)
    # skew part = central convection: +-beta_d h / 2 on face neighbours only
    assert np.isclose(np.abs(K).max(), 10.0 * h / 2)
    assert np.allclose(S_[S_ != 0][np.abs(S_[S_ != 0] - 26 / 9) > 1e-12], -1 / 9)
    rptr, cind, val = S.gen_vardiff7(n, 1, 4.0, 1.5)
    A = dense(rptr, cind, val, N)
    assert np.allclose(A, A.T)
    assert np.all(np.linalg.eigvalsh(A) > 0)
    rptr, cind, val = S.gen_unstructured(2000, 1)
    for i in range(0, 2000, 97):
        row = slice(rptr[i], rptr[i + 1])
        cols, vals = cind[row], val[row]
        off = np.abs(vals[cols != i]).sum()
        d = vals[cols == i][0]
        assert d == np.float32(1.1 * off) if off > 0 else d == 1.0
        assert np.all(vals.astype(np.float32) == vals)
        assert np.all(np.diff(cols) > 0)
    rptr, cind, val = S.gen_elasticity3(3, 2)
    N = 81
    A = dense(rptr, cind, val, N)
    assert np.allclose(A, A.T)
    assert np.all(np.linalg.eigvalsh(A) > 0)
    bv, bc, br = S.csr2bcsr(val, cind, rptr, N, 3, 3)
    assert len(bc) == 7 * 27 - 6 * 9
    x = S.gen_uniform(1, 5)
    assert np.all(np.abs(x) <= 1) and np.array_equal(x, S.gen_uniform(1, 5))
