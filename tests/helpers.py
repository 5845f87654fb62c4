"""Shared helpers for the parity tests: build systems with the library's host setup,
factor them with the oracle, and compare results."""
from __future__ import annotations

import numpy as np

import senk_amd as S
from oracle import oracle as O


def abmc_system(rptr, cind, val, bsize, method="connect", isSym=False):
    """ABMC ordering + reordering (library host code; pinned bit-exact separately)."""
    N = len(rptr) - 1
    nc, sc, LP, RP = S.abmc_permutation(cind, rptr, N, bsize, isSym, method)
    v2, c2, r2 = S.reordering(val, cind, rptr, LP, RP, N)
    return dict(N=N, nc=nc, cptr=sc, LP=LP, RP=RP, val=v2, cind=c2, rptr=r2, bsize=bsize)


def oracle_ilu0(val, cind, rptr):
    N = len(rptr) - 1
    v = O.f64(val).copy()
    st = O.orc().or_ilu0_d(O.P(v), O.P(O.i32(cind)), O.P(O.i32(rptr)), N)
    assert st == 0
    return v


def oracle_split_ldu(val, cind, rptr):
    """Split "L-DU" with inverted diagonal via the oracle; returns (L, U) as (val, cind, rptr)."""
    N = len(rptr) - 1
    nnz = int(rptr[-1])
    lv, uv = np.zeros(nnz), np.zeros(nnz)
    lc, uc = np.zeros(nnz, np.int32), np.zeros(nnz, np.int32)
    lr, ur = np.zeros(N + 1, np.int32), np.zeros(N + 1, np.int32)
    d = np.zeros(N)
    O.orc().or_split_d(O.P(O.f64(val)), O.P(O.i32(cind)), O.P(O.i32(rptr)), O.P(lv), O.P(lc), O.P(lr), O.P(uv),
                       O.P(uc), O.P(ur), O.P(d), N, 1, 1)
    return (lv[: lr[-1]].copy(), lc[: lr[-1]].copy(), lr), (uv[: ur[-1]].copy(), uc[: ur[-1]].copy(), ur)


def ilu_factors(sys_):
    f = oracle_ilu0(sys_["val"], sys_["cind"], sys_["rptr"])
    return oracle_split_ldu(f, sys_["cind"], sys_["rptr"])


def dense(rptr, cind, val, N, M=None):
    A = np.zeros((N, M or N))
    for i in range(N):
        for j in range(rptr[i], rptr[i + 1]):
            A[i, cind[j]] += val[j]
    return A


def relerr(a, b):
    a, b = np.asarray(a, float), np.asarray(b, float)
    nb = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (nb if nb > 0 else 1.0)


def mm_text(kind, seed, n=40, m=None):
    """A MatrixMarket coordinate file (text) for the ingestion tests.

    kind "general": a nonsymmetric n x m matrix with a nonzero diagonal; "symmetric": the
    lower triangle of a symmetric n x n matrix; "integer": a general matrix with integer
    values.  Entries come in shuffled order with a few explicit zeros off the diagonal,
    comment lines, a blank line between entries and values written in several notations.
    Every row keeps a nonzero entry (the diagonal in the square part); no coordinate
    repeats.
    """
    rng = np.random.default_rng(seed)
    m = n if m is None else m
    A = np.where(rng.random((n, m)) < 0.12, rng.normal(size=(n, m)), 0.0)
    if kind == "integer":
        A = np.round(A * 7)
    k = min(n, m)
    A[np.arange(k), np.arange(k)] = rng.uniform(2, 4, k)
    if kind == "integer":
        A[np.arange(k), np.arange(k)] = np.round(A[np.arange(k), np.arange(k)])
    if n > m:
        A[np.arange(m, n), rng.integers(0, m, n - m)] = 1.5  # rows beyond the square part
    if kind == "symmetric":
        A = np.tril(A) + np.tril(A, -1).T
        r, c = np.nonzero(np.tril(A))
    else:
        r, c = np.nonzero(A)
    v = A[r, c]
    # explicit zeros off the diagonal
    zr = rng.integers(1, n, 6)
    zc = np.array([rng.integers(0, i) for i in zr]) if kind == "symmetric" else rng.integers(0, m, 6)
    keep = (zr != zc) & (A[zr, np.minimum(zc, m - 1)] == 0)
    z = np.unique(np.stack([zr[keep], zc[keep]], 1), axis=0).reshape(-1, 2)
    r = np.concatenate([r, z[:, 0]])
    c = np.concatenate([c, z[:, 1]])
    v = np.concatenate([v, np.zeros(len(z))])
    perm = rng.permutation(len(v))
    r, c, v = r[perm], c[perm], v[perm]
    field = "integer" if kind == "integer" else "real"
    sym = "symmetric" if kind == "symmetric" else "general"
    out = [f"%%MatrixMarket matrix coordinate {field} {sym}", "% generated for the ingestion tests", "%"]
    out.append(f"{n} {m} {len(v)}")
    for e, (i, j, x) in enumerate(zip(r, c, v)):
        if kind == "integer":
            s = str(int(x))
        elif e % 5 == 1:
            s = f"{x:.17e}"
        elif e % 5 == 2:
            s = ("+" if x >= 0 else "") + repr(float(x))
        elif e % 5 == 3:
            s = f"{x:.6g}"
        else:
            s = repr(float(x))
        sep = "   " if e % 7 == 3 else " "
        out.append(f"{i + 1}{sep}{j + 1} {s}")
        if e == len(v) // 2:
            out.append("")
    return "\n".join(out) + "\n"


def mm_from_csr(rptr, cind, val, N, symmetric=False):
    """MatrixMarket text of a CSR matrix (values exact: repr round-trips a double); with
    symmetric=True only the lower triangle is written, as the format stores it."""
    out = [f"%%MatrixMarket matrix coordinate real {'symmetric' if symmetric else 'general'}"]
    rows = np.repeat(np.arange(N), np.diff(rptr))
    keep = cind <= rows if symmetric else np.ones(len(cind), bool)
    out.append(f"{N} {N} {int(keep.sum())}")
    out += [f"{i + 1} {j + 1} {float(v)!r}" for i, j, v in zip(rows[keep], cind[keep], val[keep])]
    return "\n".join(out) + "\n"


# inputs of the command-line workflow tests (senk_amd/senk_main): name -> (generator, symmetric file)
WORKFLOW_CASES = {
    "poisson7_14_sym": (lambda: S.gen_poisson7(14), True),
    "convdiff27_10": (lambda: S.gen_convdiff27(10, 10.0), False),
    "unstructured_1500": (lambda: S.gen_unstructured(1500, 3), False),
}


def workflow_text(name):
    gen, sym = WORKFLOW_CASES[name]
    rptr, cind, val = gen()
    return mm_from_csr(rptr, cind, val, len(rptr) - 1, sym)
