"""The C++ drop-in header's kernel surface (senk::sparse, senk::blas1, senk::blas2,
senk::test) used as a reference caller would: tests/dropin/kernels.cpp, built with g++
against libsenk_amd.so, runs every SpMV and SpTRSV form and every BLAS-1/2 routine on host
arrays for T = double and T = float, and the outputs are compared with the oracle.

Bar: bit-exact where the kernel keeps the reference's per-row (or per-element) order:
the SELL and separate-diagonal SpMV, every SpTRSV form (colored and block-Jacobi solves
equal the sequential solve on the same factors), the elementwise BLAS-1 routines,
Ggen/Grot and Trsv; the CSR and BCSR SpMV (a row's sum split over lanes) to 1e-12.  Dot and
Nrm2 are reductions whose order differs from the reference's OpenMP order (as the
reference's own results differ between thread counts): 1e-14 relative to sum |x_i y_i|
in FP64.  For T = float the oracle's float restatements are the checker where they exist
(SpMV CSR/SELL/BCSR, sequential SpTRSV); the header computes the forms without an FP32
device path in double and rounds once, which the expected values reproduce.
"""
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, oracle_ilu0, oracle_split_ldu, relerr

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]
o = O.orc()
P = O.P


def _write(path, arrays):
    with open(path, "wb") as f:
        for name, a in arrays.items():
            a = np.asarray(a)
            kind = 0 if a.dtype.kind == "i" else 1
            a = np.ascontiguousarray(a, dtype=np.int32 if kind == 0 else np.float64).ravel()
            nb = name.encode()
            f.write(np.int32(len(nb)).tobytes() + nb + np.int32(kind).tobytes() + np.int64(a.size).tobytes())
            f.write(a.tobytes())


def _read(path):
    out = {}
    raw = Path(path).read_bytes()
    p = 0
    while p < len(raw):
        ln = int(np.frombuffer(raw, np.int32, 1, p)[0])
        p += 4
        name = raw[p:p + ln].decode()
        p += ln
        kind = int(np.frombuffer(raw, np.int32, 1, p)[0])
        n = int(np.frombuffer(raw, np.int64, 1, p + 4)[0])
        p += 12
        dt = {0: np.int32, 1: np.float64, 2: np.float32}[kind]
        out[name] = np.frombuffer(raw, dt, n, p).copy()
        p += n * np.dtype(dt).itemsize
    return out


def _factors(val, cind, rptr):
    return oracle_split_ldu(oracle_ilu0(val, cind, rptr), cind, rptr)


def _inputs():
    rptr, cind, val = S.gen_elasticity3(4, 1)  # N = 192: divisible by 3 and by 64
    N = len(rptr) - 1
    rng = np.random.default_rng(17)
    A = dict(N=np.array([N]), x=rng.uniform(-1, 1, N), y0=rng.uniform(0.5, 2, N) * rng.choice([-1, 1], N),
             scalars=np.array([0.37, -1.25]), A_val=val, A_cind=cind, A_rptr=rptr)
    rows = np.repeat(np.arange(N), np.diff(rptr))
    off = cind != rows
    A["Aoff_val"], A["Aoff_cind"] = val[off], cind[off]
    A["Aoff_rptr"] = np.concatenate([[0], np.cumsum(np.bincount(rows[off], minlength=N))]).astype(np.int32)
    A["diag"] = val[~off]
    A["Ab_val"], A["Ab_cind"], A["Ab_rptr"] = S.csr2bcsr(val, cind, rptr, N, 3, 3)
    A["As_val"], A["As_cind"], A["As_wid"] = S.csr2sell(val, cind, rptr, 64, N)
    A["As_len"] = np.array([64])
    (A["L_val"], A["L_cind"], A["L_rptr"]), (A["U_val"], A["U_cind"], A["U_rptr"]) = _factors(val, cind, rptr)
    nc, sc, LP, RP = S.amc_permutation(cind, rptr, N, False)
    v2, c2, r2 = S.reordering(val, cind, rptr, LP, RP, N)
    (A["Lm_val"], A["Lm_cind"], A["Lm_rptr"]), (A["Um_val"], A["Um_cind"], A["Um_rptr"]) = _factors(v2, c2, r2)
    A["amc_cptr"], A["amc_cnum"] = sc, np.array([nc])
    sy = abmc_system(rptr, cind, val, 4)
    (A["La_val"], A["La_cind"], A["La_rptr"]), (A["Ua_val"], A["Ua_cind"], A["Ua_rptr"]) = _factors(
        sy["val"], sy["cind"], sy["rptr"])
    A["abmc_cptr"], A["abmc_cnum"], A["abmc_bsize"] = sy["cptr"], np.array([sy["nc"]]), np.array([4])
    bnum = 4
    v3, c3, r3 = S.remove_off_diagonal(val, cind, rptr, N, bnum)
    (A["Lj_val"], A["Lj_cind"], A["Lj_rptr"]), (A["Uj_val"], A["Uj_cind"], A["Uj_rptr"]) = _factors(v3, c3, r3)
    A["bj_bnum"] = np.array([bnum])
    fv, fc, fr = S.alloc_block_zero(val, cind, rptr, N, 3, 3)
    L, U = _factors(fv, fc, fr)
    A["Lb_val"], A["Lb_cind"], A["Lb_rptr"] = S.csr2bcsr(*L, N, 3, 3)
    A["Ub_val"], A["Ub_cind"], A["Ub_rptr"] = S.csr2bcsr(*U, N, 3, 3)
    sb = abmc_system(rptr, cind, val, 12)
    fv, fc, fr = S.alloc_block_zero(sb["val"], sb["cind"], sb["rptr"], N, 3, 3)
    L, U = _factors(fv, fc, fr)
    A["Lba_val"], A["Lba_cind"], A["Lba_rptr"] = S.csr2bcsr(*L, N, 3, 3)
    A["Uba_val"], A["Uba_cind"], A["Uba_rptr"] = S.csr2bcsr(*U, N, 3, 3)
    A["bcsra_cptr"], A["bcsra_cnum"], A["bcsra_bsize"] = sb["cptr"], np.array([sb["nc"]]), np.array([12])
    n, m = 31, 30
    H = np.triu(rng.uniform(-1, 1, (n, n))) + 4 * np.eye(n)
    A["H"] = np.asfortranarray(H).ravel(order="F")
    A["e"] = rng.uniform(-1, 1, n)
    A["trsv_n"], A["trsv_m"] = np.array([n]), np.array([m])
    return A


def _expected(A, f32):
    """The oracle's outputs for the driver's operations, with inputs rounded to float for
    f32 (then the float restatements where the oracle has them, else the double forms on
    the rounded inputs, rounded once)."""
    N = int(A["N"][0])
    t = np.float32 if f32 else np.float64
    R = (lambda a: np.asarray(a, np.float32).astype(np.float64)) if f32 else (lambda a: np.asarray(a, np.float64))
    x = R(A["x"])
    E = {}

    keep = []

    def arr(a, real):
        b = np.ascontiguousarray(a, real if np.asarray(a).dtype.kind == "f" else np.int32)
        keep.append(b)
        return P(b)

    def spmv(fn, names, extra=(), real=t):
        y = np.zeros(N, real)
        getattr(o, fn)(*[arr(R(A[n]) if A[n].dtype.kind == "f" else A[n], real) for n in names], arr(x, real),
                       P(y), N, *extra)
        return y

    sfx = "f" if f32 else "d"
    E["spmv"] = spmv(f"or_spmv_csr_{sfx}", ["A_val", "A_cind", "A_rptr"])
    d = np.zeros(N)
    o.or_spmv_csr_diag_d(arr(R(A["Aoff_val"]), np.float64), arr(A["Aoff_cind"], np.int32),
                         arr(A["Aoff_rptr"], np.int32), arr(R(A["diag"]), np.float64), arr(x, np.float64), P(d), N)
    E["spmv_diag"] = d.astype(t)
    y = np.zeros(N, t)
    getattr(o, f"or_spmv_bcsr_{sfx}")(arr(R(A["Ab_val"]), t), arr(A["Ab_cind"], np.int32), arr(A["Ab_rptr"], np.int32),
                                      arr(x, t), P(y), N, 3, 3)
    E["spmv_bcsr"] = y
    y = np.zeros(N, t)
    getattr(o, f"or_spmv_sell_{sfx}")(arr(R(A["As_val"]), t), arr(A["As_cind"], np.int32), arr(A["As_wid"], np.int32),
                                      64, arr(x, t), P(y), N)
    E["spmv_sell"] = y

    def tri(tag, ln, un, fl, fu, extra=()):
        yl, yu = np.zeros(N, t), np.zeros(N, t)
        getattr(o, fl)(arr(R(A[ln + "_val"]), t), arr(A[ln + "_cind"], np.int32), arr(A[ln + "_rptr"], np.int32),
                       arr(x, t), P(yl), N, *extra)
        getattr(o, fu)(arr(R(A[un + "_val"]), t), arr(A[un + "_cind"], np.int32), arr(A[un + "_rptr"], np.int32),
                       P(yl), P(yu), N, *extra)
        E[tag + "_l"], E[tag + "_u"] = yl, yu

    # colored and block-Jacobi solves equal the sequential solve on the same factors
    for tag, ln, un in (("seq", "L", "U"), ("amc", "Lm", "Um"), ("abmc", "La", "Ua"), ("bj", "Lj", "Uj")):
        tri(tag, ln, un, f"or_sptrsv_l_{sfx}", f"or_sptrsv_u_{sfx}")
    for tag, ln, un in (("bcsr", "Lb", "Ub"), ("bcsra", "Lba", "Uba")):
        yl, yu = np.zeros(N), np.zeros(N)
        o.or_sptrsv_bcsr_l_d(arr(R(A[ln + "_val"]), np.float64), arr(A[ln + "_cind"], np.int32),
                             arr(A[ln + "_rptr"], np.int32), arr(x, np.float64), P(yl), N, 3, 3)
        o.or_sptrsv_bcsr_u_d(arr(R(A[un + "_val"]), np.float64), arr(A[un + "_cind"], np.int32),
                             arr(A[un + "_rptr"], np.int32), P(yl), P(yu), N, 3, 3)
        E[tag + "_l"], E[tag + "_u"] = yl, yu  # FP64; the float run is compared to tolerance
    a, b = (float(v) for v in R(A["scalars"]))
    xt, y0t = x.astype(t), R(A["y0"]).astype(t)
    at, bt = t(a), t(b)
    if f32:  # numpy float32 arithmetic rounds every operation, as the device code (no contraction)
        E["copy"] = xt.copy()
        E["scal"] = at * xt
        E["axpy"] = y0t + at * xt
        E["axpby"] = at * xt + bt * y0t
        E["axpyz"] = at * xt + y0t
        E["hadprod"] = y0t * xt
        E["haddiv"] = y0t / xt
    else:
        w = {k: y0t.copy() for k in ("axpy", "axpby", "hadprod", "haddiv")}
        xx = xt.copy()
        o.or_axpy_d(a, P(xx), P(w["axpy"]), N)
        o.or_axpby_d(a, P(xx), b, P(w["axpby"]), N)
        o.or_hadprod_d(P(xx), P(w["hadprod"]), N)
        o.or_haddiv_d(P(xx), P(w["haddiv"]), N)
        E.update(w)
        E["axpyz"], E["copy"] = np.zeros(N), np.zeros(N)
        o.or_axpyz_d(a, P(xx), P(y0t), P(E["axpyz"]), N)
        o.or_copy_d(P(xx), P(E["copy"]), N)
        E["scal"] = xx.copy()
        o.or_scal_d(a, P(E["scal"]), N)
    xd, yd = xt.astype(np.float64), y0t.astype(np.float64)
    E["dot_nrm2"] = np.array([np.dot(xd, yd), np.sqrt(np.dot(xd, xd))])
    E["dot_scale"] = np.array([np.sum(np.abs(xd * yd)), np.sqrt(np.dot(xd, xd))])
    import ctypes as C
    c, s = C.c_double(), C.c_double()
    r = o.or_ggen_d(float(at), float(bt), C.byref(c), C.byref(s))
    ct, st = t(c.value), t(s.value)
    ga, gb = C.c_double(float(at)), C.c_double(float(bt))
    o.or_grot_d(float(ct), float(st), C.byref(ga), C.byref(gb))
    E["ggen_grot"] = np.array([r, c.value, s.value, ga.value, gb.value]).astype(t)
    n, m = int(A["trsv_n"][0]), int(A["trsv_m"][0])
    yt = np.zeros(n)
    o.or_trsv_d(arr(R(A["H"]), np.float64), arr(R(A["e"]), np.float64), P(yt), n, m)
    E["trsv"] = yt[:m].astype(t)
    return E


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_dropin_kernels_match_oracle(tmp_path):
    exe = tmp_path / "kernels"
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "dropin" / "kernels.cpp"),
                        f"-L{ROOT / 'senk_amd'}", "-lsenk_amd", f"-Wl,-rpath,{ROOT / 'senk_amd'}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    A = _inputs()
    inp, outp = tmp_path / "in.bin", tmp_path / "out.bin"
    _write(inp, A)
    run = subprocess.run([str(exe), str(inp), str(outp)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    got = _read(outp)
    exact = ["spmv_diag", "spmv_sell", "seq_l", "seq_u", "amc_l", "amc_u", "abmc_l", "abmc_u",
             "bj_l", "bj_u", "copy", "scal", "axpy", "axpby", "axpyz", "hadprod", "haddiv", "ggen_grot", "trsv"]
    for f32, sfx in ((False, "_d"), (True, "_f")):
        E = _expected(A, f32)
        for key in exact:
            g = got[key + sfx]
            assert g.dtype == (np.float32 if f32 else np.float64), key
            assert np.array_equal(g, E[key]), (key, sfx, np.max(np.abs(g.astype(float) - E[key])))
        # the CSR and BCSR products split a row's sum over several lanes on the device:
        # the plain path's bar, 1e-12 relative in FP64 (FP32: its rounding)
        for key in ("spmv", "spmv_bcsr"):
            g = got[key + sfx]
            assert g.dtype == (np.float32 if f32 else np.float64), key
            assert relerr(g.astype(np.float64), E[key].astype(np.float64)) <= (1e-6 if f32 else 1e-12), (key, sfx)
        for key in ("bcsr_l", "bcsr_u", "bcsra_l", "bcsra_u"):
            g = got[key + sfx].astype(np.float64)
            if f32:
                assert relerr(g, E[key]) <= 1e-5, (key, relerr(g, E[key]))
            else:
                assert np.array_equal(g, E[key]), key
        dn = got["dot_nrm2" + sfx].astype(np.float64)
        tol = 1e-14 if not f32 else 1e-7
        assert abs(dn[0] - E["dot_nrm2"][0]) <= tol * E["dot_scale"][0]
        assert abs(dn[1] - E["dot_nrm2"][1]) <= tol * E["dot_scale"][1]
    # senk::test::RelativeResidualError prints the reference's line
    m = re.findall(r"^# test (\S+)$", run.stdout, re.M)
    assert m, run.stdout
    want = o.or_relative_residual_d(P(A["A_val"]), P(A["A_cind"]), P(A["A_rptr"]), P(A["y0"]), P(A["x"]),
                                    int(A["N"][0]))
    assert abs(float(m[-1]) - want) <= 1e-5 * want
