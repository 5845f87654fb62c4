"""senk_relative_residual_error_noalloc (no allocation: its partial sums live in the
library's file-scope device array) against senk_relative_residual_error on the same
inputs, two calls in a row, the second on a smaller system, and against the oracle.

Bar: the residual entries are the same ordered row sums in both entry points (rows split
over 8 lanes, the solvers' residual pass); the squares are summed in a different order,
so the two agree to 1e-13 relative plus the rounding a different row-sum order could
move (the oracle, which sums each row in CSR order)."""
import ctypes as C

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O

pytestmark = pytest.mark.gpu
o = O.orc()
P = O.P


def _slack(rptr, cind, val, b, x):
    N = len(rptr) - 1
    ax = np.zeros(N)
    o.or_spmv_csr_d(P(np.abs(val)), P(cind), P(rptr), P(np.abs(x)), P(ax), N)
    w = int(np.diff(rptr).max()) + 1
    return w * np.finfo(float).eps * np.linalg.norm(ax + np.abs(b)) / np.linalg.norm(b)


def test_relres_noalloc_matches_allocating_entry_point(ctx):
    rng = np.random.default_rng(12)
    print("[no-alloc] two calls: vardiff7 24^3, then unstructured 3000", flush=True)
    for rptr, cind, val in (S.gen_vardiff7(24, 3, 8.0, 1.5), S.gen_unstructured(3000, 4)):
        N = len(rptr) - 1
        b = rng.uniform(-1, 1, N)
        for x in (rng.uniform(-1, 1, N), np.zeros(N)):
            dv, dc, dr, db, dx = (ctx.array(t) for t in (val, cind, rptr, b, x))
            want = C.c_double()
            S.call("senk_relative_residual_error", ctx.h, dv.p, dc.p, dr.p, db.p, dx.p, N, 0, C.byref(want))
            dres = ctx.zeros(1)
            S.call("senk_relative_residual_error_noalloc", ctx.h, dv.p, dc.p, dr.p, db.p, dx.p, N, dres.p)
            S.call("senk_ctx_synchronize", ctx.h)
            got = float(dres.download()[0])
            assert abs(got - want.value) <= 1e-13 * want.value, (N, got, want.value)
            ref = o.or_relative_residual_d(P(val), P(cind), P(rptr), P(b), P(x), N)
            assert abs(got - ref) <= 1e-13 * ref + _slack(rptr, cind, val, b, x), (N, got, ref)


def test_relres_noalloc_rejects_bad_arguments(ctx):
    d = ctx.zeros(4)
    assert S.lib().senk_relative_residual_error_noalloc(ctx.h, d.p, d.p, d.p, d.p, d.p, 0, d.p) == 1
    assert S.lib().senk_relative_residual_error_noalloc(ctx.h, None, d.p, d.p, d.p, d.p, 4, d.p) == 1
    assert S.lib().senk_relative_residual_error_noalloc(None, d.p, d.p, d.p, d.p, d.p, 4, d.p) == 1
