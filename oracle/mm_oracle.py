"""TEST INFRASTRUCTURE ONLY: a numpy restatement of the reference's MatrixMarket
ingestion and the CSR pre-processing of its workflow, used by tests/ as the checker.

- read_matrix_market  senk_io.hpp:41-103 (header test by substring, as the reference:
  "complex" refused, "symmetric" -> Sym, "general" -> Unsym; lines opening with '%'
  skipped; PE triples read as whitespace-separated tokens; zeros dropped on request;
  CSR sorted by row, then column)
- remove_zeros        senk_matrix.hpp:29-43
- expand              senk_matrix.hpp:335-371 (keys "L" and "U")

Pinned in tests/test_matrix_market.py against the compiled reference (oracle/_ref:
ref_read_mm, ref_expand_d, ref_remove_zeros_d) where it is built, against fixtures made
from it (tests/golden/matrix_market.npz), and against closed forms: the CSR read from a
file equals the dense matrix assembled from the file's triples, and expanding the lower
triangle of a symmetric matrix gives the matrix back.

Plain Python loops: for small files only.
"""
from __future__ import annotations

import numpy as np


def read_matrix_market(path, remove_zeros=True):
    """(val, cind, rptr, N, M, shape) with shape 0 = Sym, 1 = Unsym (senk_io.hpp:26-29)."""
    with open(path, "rb") as f:
        lines = f.read().decode().split("\n")
    head = lines[0]
    if "complex" in head:  # senk_io.hpp:53-56
        raise ValueError("header line is not valid")
    if "symmetric" in head:  # :57-59
        shape = 0
    elif "general" in head:
        shape = 1
    else:
        raise ValueError("header line is not valid")
    k = 1
    while lines[k][:1] == "%":  # :62-67
        k += 1
    N, M, PE = (int(t) for t in lines[k].split()[:3])
    tok = " ".join(lines[k + 1:]).split()
    rows, cols, vals = [], [], []
    for e in range(PE):  # :77-84
        r, c, v = int(tok[3 * e]), int(tok[3 * e + 1]), float(tok[3 * e + 2])
        if remove_zeros and v == 0:
            continue
        rows.append(r - 1)
        cols.append(c - 1)
        vals.append(v)
    rows = np.array(rows, dtype=np.int64)
    cols = np.array(cols, dtype=np.int64)
    order = np.lexsort((cols, rows))  # :89-100 (row, then column within the row)
    rptr = np.zeros(N + 1, dtype=np.int32)
    np.add.at(rptr, rows + 1, 1)
    rptr = np.cumsum(rptr).astype(np.int32)
    return (np.array(vals, dtype=np.float64)[order], cols[order].astype(np.int32), rptr, N, M, shape)


def remove_zeros(val, cind, rptr, N):
    """senk_matrix.hpp:29-43: entries equal to zero leave; row pointers follow."""
    out_v, out_c, out_r = [], [], [int(rptr[0])]
    for i in range(N):
        for j in range(rptr[i], rptr[i + 1]):
            if val[j] != 0:
                out_v.append(val[j])
                out_c.append(cind[j])
        out_r.append(out_r[0] + len(out_v))
    return np.array(out_v, dtype=np.float64), np.array(out_c, dtype=np.int32), np.array(out_r, dtype=np.int32)


def _transpose(val, cind, rptr, N):
    cols = [[] for _ in range(N)]
    for i in range(N):
        for j in range(rptr[i], rptr[i + 1]):
            cols[cind[j]].append((i, val[j]))
    return cols


def expand(val, cind, rptr, N, key="L"):
    """senk_matrix.hpp:335-371.  "L": row i = stored row i, then column i of the stored
    matrix without its first entry; "U": column i, then row i without its first entry."""
    if key not in ("L", "U"):
        raise ValueError("Expand: keyword is not valid")
    cols = _transpose(val, cind, rptr, N)
    rows = [[(cind[j], val[j]) for j in range(rptr[i], rptr[i + 1])] for i in range(N)]
    first, second = (rows, cols) if key == "L" else (cols, rows)
    ov, oc, orp = [], [], [0]
    for i in range(N):
        for c, v in first[i] + second[i][1:]:
            oc.append(c)
            ov.append(v)
        orp.append(len(oc))
    return np.array(ov, dtype=np.float64), np.array(oc, dtype=np.int32), np.array(orp, dtype=np.int32)
