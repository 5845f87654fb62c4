"""MatrixMarket ingestion and the CSR pre-processing of the reference's workflow
(io::ReadMatrixMarket senk_io.hpp:41-103; RemoveZeros, Expand, Duplicate
senk_matrix.hpp:29-43, 335-389), CPU only.

The library's reader and helpers against the reference's own outputs (fixtures in
tests/golden/matrix_market.npz, written from the compiled reference by
tests/golden/make_golden.py; and the compiled reference itself where oracle/_ref is
built), against the numpy restatement in oracle/mm_oracle.py, and against closed forms.
Integer arrays must be identical and values identical bit for bit.
"""
import ctypes as C
from pathlib import Path

import numpy as np
import pytest

import senk_amd as S
from oracle import mm_oracle as MO
from oracle import oracle as O
from tests.golden.make_golden import MM_CASES
from tests.helpers import dense, mm_text

GOLDEN = np.load(Path(__file__).resolve().parent / "golden" / "matrix_market.npz", allow_pickle=False)


def _write(tmp_path, name, text):
    p = tmp_path / f"{name}.mtx"
    p.write_bytes(text.encode() if isinstance(text, str) else text)
    return p


def _same(a, b):
    a, b = np.asarray(a), np.asarray(b)
    if a.dtype == np.float64:
        return a.shape == b.shape and np.array_equal(a.view(np.int64), b.view(np.int64))
    return a.shape == b.shape and np.array_equal(a, b)


def _check_csr(got, v, c, r):
    assert _same(got[0], v) and _same(got[1], c) and _same(got[2], r)


@pytest.mark.parametrize("name", list(MM_CASES))
@pytest.mark.parametrize("rz", [1, 0])
def test_reader_matches_reference_fixture(tmp_path, name, rz):
    path = _write(tmp_path, name, GOLDEN[f"{name}_text"].tobytes())
    g = {k: GOLDEN[f"{name}_rz{rz}_{k}"] for k in ("val", "cind", "rptr", "dims")}
    got = S.read_matrix_market(path, rz)
    _check_csr(got, g["val"], g["cind"], g["rptr"])
    assert (got[3], got[4], got[5]) == tuple(int(x) for x in g["dims"])
    orc = MO.read_matrix_market(path, rz)
    _check_csr(orc, g["val"], g["cind"], g["rptr"])
    assert orc[3:] == got[3:]


@pytest.mark.parametrize("name", list(MM_CASES))
def test_remove_zeros_and_duplicate_match_reference_fixture(tmp_path, name):
    path = _write(tmp_path, name, GOLDEN[f"{name}_text"].tobytes())
    v, c, r, N, M, _ = S.read_matrix_market(path, 0)
    g = [GOLDEN[f"{name}_rmz_{k}"] for k in ("val", "cind", "rptr")]
    _check_csr(S.remove_zeros(v, c, r, N), *g)
    _check_csr(MO.remove_zeros(v, c, r, N), *g)
    # removing zeros after the read equals reading with removal
    _check_csr(S.remove_zeros(v, c, r, N), *S.read_matrix_market(path, 1)[:3])
    _check_csr(S.duplicate(v, c, r, N), v, c, r)


@pytest.mark.parametrize("key", ["L", "U"])
def test_expand_matches_reference_fixture(tmp_path, key):
    path = _write(tmp_path, "symmetric", GOLDEN["symmetric_text"].tobytes())
    v, c, r, N, M, shape = S.read_matrix_market(path, 1)
    assert shape == S.SENK_SHAPE_SYM and N == M
    if key == "U":
        v, c, r = S.csr2csc(v, c, r, N, N)
    g = [GOLDEN[f"symmetric_expand{key}_{k}"] for k in ("val", "cind", "rptr")]
    _check_csr(S.expand(v, c, r, N, key), *g)
    _check_csr(MO.expand(v, c, r, N, key), *g)


def _triples(text):
    """(rows, cols, vals, N, M, symmetric) straight from the text, for the closed forms."""
    lines = [ln for ln in text.splitlines() if ln.strip() and not ln.startswith("%")]
    N, M, PE = (int(t) for t in lines[0].split())
    tok = " ".join(lines[1:]).split()
    r = np.array([int(tok[3 * e]) - 1 for e in range(PE)])
    c = np.array([int(tok[3 * e + 1]) - 1 for e in range(PE)])
    v = np.array([float(tok[3 * e + 2]) for e in range(PE)])
    return r, c, v, N, M, "symmetric" in text.splitlines()[0]


@pytest.mark.parametrize("kind,n,m", [("general", 45, None), ("general", 20, 37), ("general", 37, 20),
                                      ("symmetric", 50, None), ("integer", 41, None)])
@pytest.mark.parametrize("seed", [1, 2])
def test_reader_closed_forms(tmp_path, kind, n, m, seed):
    text = mm_text(kind, seed, n, m)
    path = _write(tmp_path, "a", text)
    r, c, v, N, M, sym = _triples(text)
    D = np.zeros((N, M))
    D[r, c] = v
    for rz in (1, 0):
        val, cind, rptr, n_, m_, shape = S.read_matrix_market(path, rz)
        assert (n_, m_) == (N, M) and shape == (S.SENK_SHAPE_SYM if sym else S.SENK_SHAPE_UNSYM)
        assert np.array_equal(dense(rptr, cind, val, N, M), D)
        assert len(val) == (np.count_nonzero(v) if rz else len(v))
        assert not rz or np.all(val != 0)
        for i in range(N):  # columns strictly ascending in each row
            assert np.all(np.diff(cind[rptr[i]:rptr[i + 1]]) > 0)
        _check_csr((val, cind, rptr), *MO.read_matrix_market(path, rz)[:3])
    if sym:
        # the stored lower triangle expands to the symmetric matrix ("L"), and its
        # transpose (the upper triangle) expands to the same matrix ("U")
        val, cind, rptr = S.read_matrix_market(path, 1)[:3]
        full = D + np.tril(D, -1).T
        assert S.check_structure(cind, rptr, N)
        ev, ec, er = S.expand(val, cind, rptr, N, "L")
        assert len(ev) == 2 * len(val) - N
        assert np.array_equal(dense(er, ec, ev, N), full)
        uv, uc, ur = S.csr2csc(val, cind, rptr, N, N)
        fv, fc, fr = S.expand(uv, uc, ur, N, "U")
        assert np.array_equal(dense(fr, fc, fv, N), full)
        for i in range(N):
            assert np.all(np.diff(ec[er[i]:er[i + 1]]) > 0) and np.all(np.diff(fc[fr[i]:fr[i + 1]]) > 0)


def test_reader_against_compiled_reference(tmp_path):
    ref = O.ref()
    if ref is None:
        pytest.skip("oracle/_ref not built (reference tree absent)")
    P = O.P
    for seed in range(6):
        for kind, n, m in (("general", 60, None), ("general", 30, 47), ("symmetric", 64, None), ("integer", 50, 35)):
            path = _write(tmp_path, f"{kind}{seed}", mm_text(kind, 100 + seed, n, m))
            for rz in (1, 0):
                cap = 8192
                v, c, r, d = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(3, np.int32)
                nnz = ref.ref_read_mm(str(path).encode(), rz, P(v), P(c), P(r), P(d), cap, cap)
                assert nnz >= 0
                got = S.read_matrix_market(path, rz)
                _check_csr(got, v[:nnz], c[:nnz], r[:d[0] + 1])
                assert got[3:] == tuple(int(x) for x in d)
                if kind == "symmetric" and rz:
                    N = got[3]
                    ev, ec, er = np.zeros(2 * cap), np.zeros(2 * cap, np.int32), np.zeros(N + 1, np.int32)
                    k = ref.ref_expand_d(P(got[0]), P(got[1]), P(got[2]), N, b"L", P(ev), P(ec), P(er), 2 * cap)
                    _check_csr(S.expand(*got[:3], N, "L"), ev[:k], ec[:k], er)


def test_oracle_plants_are_caught(tmp_path):
    """The fixture comparison fails for a reader that keeps rows unsorted or keeps zeros."""
    path = _write(tmp_path, "general", GOLDEN["general_text"].tobytes())
    v, c, r, N, M, _ = MO.read_matrix_market(path, 0)
    g = [GOLDEN[f"general_rz1_{k}"] for k in ("val", "cind", "rptr")]
    assert not (_same(v, g[0]) and _same(c, g[1]) and _same(r, g[2]))  # zeros kept
    v, c, r = MO.read_matrix_market(path, 1)[:3]
    i = int(np.argmax(np.diff(r) >= 2))
    assert r[i + 1] - r[i] >= 2
    c2 = c.copy()
    c2[r[i]:r[i + 1]] = c2[r[i]:r[i + 1]][::-1]  # a row left in file order
    assert _same(c, g[1]) and not _same(c2, g[1])


BAD = {
    "complex": "%%MatrixMarket matrix coordinate complex general\n2 2 1\n1 1 1.0 0.0\n",
    "pattern": "%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n",
    "hermitian": "%%MatrixMarket matrix coordinate real hermitian\n2 2 1\n1 1 1.0\n",
    "skew": "%%MatrixMarket matrix coordinate real skew-symmetric\n2 2 1\n2 1 1.0\n",
    "array": "%%MatrixMarket matrix array real general\n2 2\n1.0\n0.0\n0.0\n1.0\n",
    "no_banner": "2 2 1\n1 1 1.0\n",
    "size_line": "%%MatrixMarket matrix coordinate real general\n2 x 1\n1 1 1.0\n",
    "short": "%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n2 2 1.0\n",
    "outside": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n3 2 1.0\n",
    "zero_index": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n0 2 1.0\n",
    "bad_value": "%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n2 2 x\n",
}


@pytest.mark.parametrize("name", list(BAD))
def test_reader_refuses_what_it_does_not_read(tmp_path, name):
    path = _write(tmp_path, name, BAD[name])
    out = [C.c_void_p() for _ in range(3)] + [C.c_int() for _ in range(3)]
    st = S.lib().senk_read_matrix_market(str(path).encode(), *(C.byref(o) for o in out), 1)
    assert st == 13, S.STATUS.get(st, st)  # SENK_ERR_FORMAT
    assert S.lib().senk_last_error()


def test_reader_missing_file_and_arguments(tmp_path):
    out = [C.c_void_p() for _ in range(3)] + [C.c_int() for _ in range(3)]
    st = S.lib().senk_read_matrix_market(str(tmp_path / "absent.mtx").encode(), *(C.byref(o) for o in out), 1)
    assert st == 12  # SENK_ERR_IO
    with pytest.raises(S.SenkError, match="SENK_ERR_IO"):
        S.read_matrix_market(tmp_path / "absent.mtx")
    assert S.lib().senk_read_matrix_market(None, *(C.byref(o) for o in out), 1) == 1
    rptr, cind, val = S.gen_poisson7(3)
    with pytest.raises(S.SenkError, match="SENK_ERR_ARG"):
        S.expand(val, cind, rptr, 27, "X")


def test_reader_edge_cases(tmp_path):
    # case-insensitive banner, CRLF line ends, comments and blank lines before the size
    # line, an empty matrix, and rows without entries in the middle and at the end
    text = ("%%matrixmarket MATRIX Coordinate REAL General\r\n% c\r\n\r\n%\r\n4 3 3\r\n"
            "3 1 -2.5e0\r\n1 3 +4\r\n3 3 0\r\n")
    val, cind, rptr, N, M, shape = S.read_matrix_market(_write(tmp_path, "e", text), 1)
    assert (N, M, shape) == (4, 3, S.SENK_SHAPE_UNSYM)
    assert rptr.tolist() == [0, 1, 1, 2, 2] and cind.tolist() == [2, 0] and val.tolist() == [4.0, -2.5]
    val, cind, rptr, N, M, _ = S.read_matrix_market(_write(tmp_path, "e", text), 0)
    assert rptr.tolist() == [0, 1, 1, 3, 3] and cind.tolist() == [2, 0, 2]
    empty = "%%MatrixMarket matrix coordinate real general\n3 3 0\n"
    val, cind, rptr, N, M, _ = S.read_matrix_market(_write(tmp_path, "z", empty))
    assert N == 3 and len(val) == 0 and rptr.tolist() == [0, 0, 0, 0]
    # the values are the correctly rounded parses of the file's text
    vals = ["0.1", "1e-300", "-123456789.123456789", "6.02214076e23", "2.2250738585072014e-308"]
    text = "%%MatrixMarket matrix coordinate real general\n5 5 5\n" + "".join(
        f"{i + 1} {i + 1} {s}\n" for i, s in enumerate(vals))
    val = S.read_matrix_market(_write(tmp_path, "v", text))[0]
    assert _same(val, np.array([float(s) for s in vals]))


def test_size_query(tmp_path):
    for name in MM_CASES:
        text = GOLDEN[f"{name}_text"].tobytes()
        path = _write(tmp_path, name, text)
        N, M, E, sh = C.c_int(), C.c_int(), C.c_longlong(), C.c_int()
        S.call("senk_matrix_market_size", str(path).encode(), C.byref(N), C.byref(M), C.byref(E), C.byref(sh))
        size_line = [ln for ln in text.decode().splitlines()[1:] if ln.strip() and not ln.startswith("%")][0]
        assert [N.value, M.value, E.value] == [int(t) for t in size_line.split()]
        assert sh.value == int(GOLDEN[f"{name}_rz1_dims"][2])
    # a comment line longer than the reader's line buffer, then the size line
    text = "%%MatrixMarket matrix coordinate real general\n%" + "x" * 10000 + "\n\n7 5 1\n1 1 2.0\n"
    path = _write(tmp_path, "long", text)
    S.call("senk_matrix_market_size", str(path).encode(), C.byref(N), C.byref(M), C.byref(E), C.byref(sh))
    assert (N.value, M.value, E.value, sh.value) == (7, 5, 1, S.SENK_SHAPE_UNSYM)
    assert S.read_matrix_market(path)[3:5] == (7, 5)
    assert S.lib().senk_matrix_market_size(str(tmp_path / "absent").encode(), C.byref(N), C.byref(M), C.byref(E),
                                           C.byref(sh)) == 12


def test_large_file_parallel_parse(tmp_path):
    """A file above the reader's parallel threshold (1 MB) whose triples are split across
    lines and runs of blank space, so that chunk cuts fall inside triples: the same arrays
    as the oracle (and the reference where built); a bad entry near the end is reported
    with its index in file order."""
    rng = np.random.default_rng(9)
    n, k = 5000, 120000
    r = rng.integers(1, n + 1, k)
    c = rng.integers(1, n + 1, k)
    key = np.unique(r.astype(np.int64) * (n + 1) + c, return_index=True)[1]
    r, c = r[np.sort(key)], c[np.sort(key)]
    v = np.round(rng.normal(size=len(r)), 6)
    v[::97] = 0.0
    seps = ["\n", " ", "\t", "\n\n", "  \n "]
    body = "".join(f"{i}{seps[e % 5]}{j}{seps[(e + 2) % 5]}{x!r}\n"
                   for e, (i, j, x) in enumerate(zip(r.tolist(), c.tolist(), v.tolist())))
    text = f"%%MatrixMarket matrix coordinate real general\n{n} {n} {len(r)}\n" + body
    assert len(text) > (1 << 20)
    path = _write(tmp_path, "big", text)
    for rz in (1, 0):
        got = S.read_matrix_market(path, rz)
        _check_csr(got, *MO.read_matrix_market(path, rz)[:3])
    ref = O.ref()
    if ref is not None:
        cap = len(r) + 8
        rv, rc, rr, d = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(n + 1, np.int32), np.zeros(3, np.int32)
        nnz = ref.ref_read_mm(str(path).encode(), 1, O.P(rv), O.P(rc), O.P(rr), O.P(d), cap, n + 1)
        _check_csr(S.read_matrix_market(path, 1), rv[:nnz], rc[:nnz], rr)
    lines = text.split("\n")
    bad = "\n".join(lines[:-40]) + "\n" + f"{n + 1} 1 1.0\n" + "\n".join(lines[-40:])
    bad = bad.replace(f"{n} {n} {len(r)}", f"{n} {n} {len(r) + 1}", 1)
    out = [C.c_void_p() for _ in range(3)] + [C.c_int() for _ in range(3)]
    st = S.lib().senk_read_matrix_market(str(_write(tmp_path, "bad", bad)).encode(), *(C.byref(o) for o in out), 1)
    assert st == 13 and b"lies outside the matrix" in S.lib().senk_last_error()
