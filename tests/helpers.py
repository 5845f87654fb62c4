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
