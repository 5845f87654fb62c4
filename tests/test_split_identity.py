"""The algebra of the split (Eisenstat) form of the mixed inner BiCGStab, on the CPU.

The library takes an ILU(0) as "D-ILU" when no elimination step touches an
off-diagonal entry (a triangle-free matrix graph).  Then, with the oracle's IKJ ILU(0)
(unit lower L, upper U) and D = diag(U), E = 2 D - diag(A):
  * the factors' off-diagonal entries are A's own (u_ij = a_ij, l_ij d_j = a_ij);
  * A = L D + U - E;
  * L^-1 A U^-1 v = D t + L^-1 (v - E t) with t = U^-1 v, the product the device forms
    from one U solve, one L solve and two vector passes.
On a 27-point stencil (triangles in the graph) the first two fail, which is why the
library checks the pattern before it uses the split form.
"""
import numpy as np

import senk_amd as S
from tests.helpers import abmc_system, oracle_ilu0


def _dense_factors(sy):
    N = sy["N"]
    F = oracle_ilu0(sy["val"], sy["cind"], sy["rptr"])
    A = np.zeros((N, N))
    LU = np.zeros((N, N))
    for i in range(N):
        for j in range(sy["rptr"][i], sy["rptr"][i + 1]):
            A[i, sy["cind"][j]] = sy["val"][j]
            LU[i, sy["cind"][j]] = F[j]
    L = np.tril(LU, -1) + np.eye(N)
    U = np.triu(LU)
    return A, L, U


def test_seven_point_ilu0_is_dilu_and_split_product_holds():
    rptr, cind, val = S.gen_vardiff7(6, 3, 8.0, 1.5)
    sy = abmc_system(rptr, cind, val, 4)
    A, L, U = _dense_factors(sy)
    off = A != 0
    np.fill_diagonal(off, False)
    up = np.triu(off, 1)
    lo = np.tril(off, -1)
    # upper off-diagonal entries untouched, lower ones a_ij / d_j
    assert np.array_equal(U[up], A[up])
    d = np.diag(U).copy()
    LD = L * d[None, :]
    assert np.allclose(LD[lo], A[lo], rtol=1e-14, atol=0)
    E = 2 * d - np.diag(A)
    assert np.abs(LD + U - np.diag(E) - A).max() <= 1e-12 * np.abs(A).max()
    # the pivots do change (not plain Gauss-Seidel)
    assert np.abs(d - np.diag(A)).max() > 1e-3 * np.abs(np.diag(A)).max()
    rng = np.random.default_rng(0)
    for _ in range(3):
        v = rng.uniform(-1, 1, sy["N"])
        want = np.linalg.solve(L, A @ np.linalg.solve(U, v))
        t = np.linalg.solve(U, v)
        got = d * t + np.linalg.solve(L, v - E * t)
        assert np.linalg.norm(got - want) <= 1e-12 * np.linalg.norm(want)


def test_twenty_seven_point_ilu0_is_not_dilu():
    rptr, cind, val = S.gen_convdiff27(6, 10.0)
    sy = abmc_system(rptr, cind, val, 4)
    A, L, U = _dense_factors(sy)
    d = np.diag(U).copy()
    E = 2 * d - np.diag(A)
    err = np.abs(L * d[None, :] + U - np.diag(E) - A).max()
    assert err > 1e-3 * np.abs(A).max(), err
