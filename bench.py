"""Benchmark of SenK's preconditioned Krylov hot path on one MI355X.

Headline workload (one GPU): the 3D 7-point variable-coefficient diffusion system
-div(k grad u) = f on a 128^3 grid (n = 2,097,152, nnz = 14,581,760; k = exp(Gaussian
random field, correlation length 8 cells, sigma 1.5), seed 1; harmonic-mean face
coefficients), ILU(0), right-hand side b = A x* with x* ~ U[-1,1] from seed 1, x0 = 0,
target 1e-10 relative residual, under three ABMC orderings: "connect" blocks of 8 (_b8)
and of 4 (_b4), and "simple" blocks of 8 consecutive rows (_s8: the block graph of the
7-point grid is 2-colorable, so each triangular solve has 2 levels).  One step is one full
solve with every input resident in HBM (setup timed separately, as the reference's
driver times only the solve).

  value          time-to-1e-10 of the fastest mixed-precision solve: FP64 flexible GCR
                 outer iteration around FP32 inner Krylov solves (BiCGStab, or CG: the
                 system is symmetric positive definite) with FP32 ILU(0) factors, in the
                 split form where the factorization allows it (ms per solve, median of
                 --steps); config.mixed names the solve
  fp64_gmres30   the plain FP64 ILU(0)-GMRES(30) solve it is compared with (the
                 reference's AbmcIluGmresm algorithm, modified Gram-Schmidt), at its
                 fastest ordering; fp64_gmres30_cgs2 the same with two-pass classical
                 Gram-Schmidt (reported beside it, not the baseline)
  fp64_ic0_cg    FP64 conjugate gradients with the incomplete Cholesky factor (the system
                 is symmetric positive definite); mixed_cg_* the nested mixed form with an
                 FP32 inner CG
  learned_*      with the trained message-passing correction: learned_fgcr30 /
                 learned_mixed_fgcr correct every ILU application of an FGCR (weights
                 senk_amd/data/gnn_vardiff.npz); learned_mixed_cg corrects the inner
                 CG's result once per outer FGCR step of the nested solve (weights
                 senk_amd/data/gnn_vardiff_nested.npz, trained on that solve's pairs)
  plain_fp64_c2  the 27-point convection-diffusion 128^3 system (Pe 10, ABMC blocks of
                 16): FP64 ILU(0)-GMRES(30) and ILU(0)-BiCGStab, and SpMV rates
  roofline       FP32 ILU(0) application of the headline solve's system (L and U sweeps,
                 the dominant work of an inner step): algorithmic bytes / average
                 application time (HIP events on the library stream, back-to-back) against
                 8 TB/s; roofline_spmv the same for the FP32 SELL-C-sigma SpMV; *_cold the
                 same for single launches after a 1 GiB stream (nothing left in the 256 MB
                 last-level cache)
  cpu_baseline   the reference's own OpenMP AbmcIluGmresm (oracle/_ref) on the host cores,
                 the full FP64 solve of the same system and factors
  ingest         MatrixMarket reading + symmetric expansion on the host: the library's
                 reader against the reference's io::ReadMatrixMarket + Expand (64^3 file)

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128] [--kernels-only]
       [--sweep-modes] [--learned-only] [--run NAME] [--ingest] [--no-cpu] [--no-c2]
       [--no-ingest]
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import statistics
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

import senk_amd as S  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "time-to-1e-10 residual & iters: mixed/learned vs FP64 ILU(0)-GMRES; SpMV GB/s"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, sv, si=4):
    return nnz * (sv + si) + (n + 1) * si + 2 * n * sv


def ilu_bytes(n, lnnz, unnz, sv, si=4):
    return (lnnz + unnz) * (sv + si) + 2 * (n + 1) * si + 4 * n * sv


def build_system_kind(kind, n, bsize, seed=1, method="connect"):
    """kind "convdiff": 27-point convection-diffusion on n^3, Pe 10; "vardiff": 7-point
    -div(k grad u) on n^3 with k = exp(Gaussian random field, correlation length 8, sigma
    1.5); "unstructured": n rows, lognormal row lengths, banded-random columns, diagonal 1.1x
    the row's absolute sum; "elasticity": 3 dof per node on n^3, 3x3-block 7-point stencil,
    nu 0.3, lognormal moduli."""
    t0 = time.perf_counter()
    if kind == "convdiff":
        rptr, cind, val = S.gen_convdiff27(n, 10.0)
    elif kind == "vardiff":
        rptr, cind, val = S.gen_vardiff7(n, seed, 8.0, 1.5)
    elif kind == "unstructured":
        rptr, cind, val = S.gen_unstructured(n, seed)
    elif kind == "elasticity":
        rptr, cind, val = S.gen_elasticity3(n, seed)
    else:
        raise ValueError(kind)
    N = len(rptr) - 1
    xs = S.gen_uniform(1, N)
    t1 = time.perf_counter()
    nc, cptr, LP, RP = S.abmc_permutation(cind, rptr, N, bsize, False, method)
    t2 = time.perf_counter()
    val2, cind2, rptr2 = S.reordering(val, cind, rptr, LP, RP, N)
    t3 = time.perf_counter()
    rows = np.repeat(np.arange(N), np.diff(rptr))
    b = np.bincount(rows, weights=val * xs[cind], minlength=N)
    return dict(N=N, nnz=int(rptr[-1]), rptr=rptr2, cind=cind2, val=val2, b=b[LP].copy(), xs=xs[LP].copy(), nc=nc,
                cptr=cptr, bsize=bsize, method=method, t_gen=t1 - t0, t_abmc=t2 - t1, t_reorder=t3 - t2)


class Device:
    """A system on the device: matrix (FP64 + FP32 values, SELL-C-sigma) and ILU(0) (FP64 + FP32);
    kind=S.SENK_IC0 for the incomplete Cholesky factor instead."""

    def __init__(self, ctx, sy, kind=S.SENK_ILU0, p=0, bnl=0, bnw=0):
        self.ctx, self.sy, self.N = ctx, sy, sy["N"]
        self.A = C.c_void_p()
        t0 = time.perf_counter()
        S.call("senk_mat_create", ctx.h, self.N, S._ptr(sy["rptr"]), S._ptr(sy["cind"]), S._ptr(sy["val"]), 1 | 2,
               C.byref(self.A))
        self.t_upload = time.perf_counter() - t0
        self.M = C.c_void_p()
        opts = S.senk_ilu_opts(kind, p, bnl, bnw, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32), 0)
        t0 = time.perf_counter()
        S.call("senk_prec_create_ilu", ctx.h, self.A, S._ptr(sy["rptr"]), S._ptr(sy["cind"]), S._ptr(sy["val"]),
               S._ptr(sy["cptr"]), sy["nc"], sy["bsize"], C.byref(opts), C.byref(self.M))
        self.t_ilu = time.perf_counter() - t0
        self.db = ctx.array(sy["b"])
        ln, un = C.c_int(), C.c_int()
        S.call("senk_prec_factor_nnz", self.M, C.byref(ln), C.byref(un), None, None)
        self.lnnz, self.unnz = ln.value, un.value

    def solve(self, keep_x=False, **kw):
        dx = self.ctx.zeros(self.N)
        o = S.solve_opts(**kw)
        res = S.senk_result()
        S.call("senk_solve", self.ctx.h, self.A, self.M, self.db.p, dx.p, C.byref(o), C.byref(res), None, 0)
        d = res.as_dict()
        x = dx.download()
        d["err_vs_xstar"] = float(np.linalg.norm(x - self.sy["xs"]) / np.linalg.norm(self.sy["xs"]))
        if keep_x:
            d["x"] = x
        return d

    def factors(self):
        """The device factors (CSR: L strict lower, U with the inverted diagonal first)."""
        N = self.N
        lr, lc, lv = np.zeros(N + 1, np.int32), np.zeros(self.lnnz, np.int32), np.zeros(self.lnnz)
        ur, uc, uv = np.zeros(N + 1, np.int32), np.zeros(self.unnz, np.int32), np.zeros(self.unnz)
        S.call("senk_prec_download", self.M, S._ptr(lr), S._ptr(lc), S._ptr(lv), S._ptr(ur), S._ptr(uc), S._ptr(uv))
        return (lv, lc, lr), (uv, uc, ur)

    def close(self):
        S.lib().senk_prec_destroy(self.M)
        S.lib().senk_mat_destroy(self.A)


def time_kernel(ctx, fn, reps):
    """Average time of fn() in ms, HIP events on the library's stream."""
    e0, e1 = C.c_void_p(), C.c_void_p()
    S.call("senk_event_create", ctx.h, C.byref(e0))
    S.call("senk_event_create", ctx.h, C.byref(e1))
    fn()
    S.call("senk_event_record", ctx.h, e0)
    for _ in range(reps):
        fn()
    S.call("senk_event_record", ctx.h, e1)
    ms = C.c_float()
    S.call("senk_event_elapsed_ms", e0, e1, C.byref(ms))
    S.lib().senk_event_destroy(e0)
    S.lib().senk_event_destroy(e1)
    return ms.value / reps


def time_cold(ctx, fn, reps, flush):
    """Average time of fn() in ms, each call timed alone with HIP events on the library's
    stream right after flush(), which reads 1 GiB: none of fn's data is then left in the
    256 MB last-level cache, so every byte comes from HBM."""
    e0, e1 = C.c_void_p(), C.c_void_p()
    S.call("senk_event_create", ctx.h, C.byref(e0))
    S.call("senk_event_create", ctx.h, C.byref(e1))
    fn()
    total = 0.0
    for _ in range(reps):
        flush()
        S.call("senk_event_record", ctx.h, e0)
        fn()
        S.call("senk_event_record", ctx.h, e1)
        ms = C.c_float()
        S.call("senk_event_elapsed_ms", e0, e1, C.byref(ms))
        total += ms.value
    S.lib().senk_event_destroy(e0)
    S.lib().senk_event_destroy(e1)
    return total / reps


def kernel_rates(ctx, dev, reps=50, sweep_modes=False, cold=True):
    """Back-to-back launches (warm: the 128^3 7-point matrix, 142 MB in FP32, stays largely
    in the 256 MB last-level cache from one launch to the next, as it does between the
    products of a solve) and, with cold, single launches after a 1 GiB stream."""
    N, nnz = dev.N, dev.sy["nnz"]
    dx = ctx.array(np.random.default_rng(0).uniform(-1, 1, N))
    dy = ctx.zeros(N)
    dx32, dy32 = ctx.zeros(N, np.float32), ctx.zeros(N, np.float32)
    S.call("senk_convert_f64_f32", ctx.h, dx.p, dx32.p, N)
    t64 = time_kernel(ctx, lambda: S.call("senk_mat_spmv_f64", dev.A, dx.p, dy.p), reps)
    t32 = time_kernel(ctx, lambda: S.call("senk_mat_spmv_f32", dev.A, dx32.p, dy32.p), reps)
    p64 = time_kernel(ctx, lambda: S.call("senk_prec_apply_f64", dev.M, dx.p, dy.p), 20)
    p32 = time_kernel(ctx, lambda: S.call("senk_prec_apply_f32", dev.M, dx32.p, dy32.p), 20)
    b64, b32 = spmv_bytes(N, nnz, 8), spmv_bytes(N, nnz, 4)
    i64, i32 = ilu_bytes(N, dev.lnnz, dev.unnz, 8), ilu_bytes(N, dev.lnnz, dev.unnz, 4)
    out = {"spmv_sell_f64_us": t64 * 1e3, "spmv_sell_f64_GBps": b64 / t64 / 1e6,
           "spmv_sell_f32_us": t32 * 1e3, "spmv_sell_f32_GBps": b32 / t32 / 1e6,
           "ilu_apply_f64_us": p64 * 1e3, "ilu_apply_f64_GBps": i64 / p64 / 1e6,
           "ilu_apply_f32_us": p32 * 1e3, "ilu_apply_f32_GBps": i32 / p32 / 1e6,
           "colors": dev.sy["nc"], "bsize": dev.sy["bsize"]}
    if cold:
        n = (1 << 29) // 8  # two vectors of 512 MiB, read by a dot product (no dirty lines)
        fa, fb, fr = ctx.zeros(n), ctx.zeros(n), ctx.zeros(1)
        flush = lambda: S.call("senk_dot_f64_dev", ctx.h, fa.p, fb.p, n, fr.p)
        c64 = time_cold(ctx, lambda: S.call("senk_mat_spmv_f64", dev.A, dx.p, dy.p), 10, flush)
        c32 = time_cold(ctx, lambda: S.call("senk_mat_spmv_f32", dev.A, dx32.p, dy32.p), 10, flush)
        q32 = time_cold(ctx, lambda: S.call("senk_prec_apply_f32", dev.M, dx32.p, dy32.p), 10, flush)
        out["cold"] = {"spmv_sell_f64_us": c64 * 1e3, "spmv_sell_f64_GBps": b64 / c64 / 1e6,
                       "spmv_sell_f32_us": c32 * 1e3, "spmv_sell_f32_GBps": b32 / c32 / 1e6,
                       "ilu_apply_f32_us": q32 * 1e3, "ilu_apply_f32_GBps": i32 / q32 / 1e6}
        del fa, fb, fr
    if not sweep_modes:
        return out
    # triangular-solve launch modes: 1 one cooperative sweep per triangle, 3 one launch per
    # triangle with a ticket stream per XCD (the default), 2 one ticketed sweep per
    # triangle, 0 one launch per level; the default is restored after
    for mode in (1, 3, 2, 0):
        S.call("senk_prec_set_trsv_mode", dev.M, mode)
        out[f"ilu_apply_f64_us_mode{mode}"] = time_kernel(
            ctx, lambda: S.call("senk_prec_apply_f64", dev.M, dx.p, dy.p), 20) * 1e3
        out[f"ilu_apply_f32_us_mode{mode}"] = time_kernel(
            ctx, lambda: S.call("senk_prec_apply_f32", dev.M, dx32.p, dy32.p), 20) * 1e3
    S.call("senk_prec_set_trsv_mode", dev.M, -1)
    S.call("senk_prec_check", dev.M)
    return out


def copy_rate(ctx, nbytes=1 << 30, reps=10):
    """A plain device copy of known size (the counter calibration of the roofline traffic)."""
    n = nbytes // 8
    a, b = ctx.zeros(n), ctx.zeros(n)
    t = time_kernel(ctx, lambda: S.call("senk_copy_f64", ctx.h, a.p, b.p, n), reps)
    return {"copy_bytes": 2 * n * 8, "copy_us": t * 1e3, "copy_GBps": 2 * n * 8 / t / 1e6}


def traffic_from_profile(name):
    """HBM bytes per launch (per application for the ILU sweeps) from a committed counter
    summary under profiles/: FETCH_SIZE and WRITE_SIZE (KiB per dispatch) scaled by what
    the same counters read for a device copy of known size."""
    path = ROOT / "profiles" / name
    try:
        vals = {}
        for line in path.read_text().splitlines():
            if line.startswith("traffic:"):  # "traffic: key=value key=value ..."
                for item in line.split(":", 1)[1].split():
                    k, v = item.split("=")
                    vals[k] = float(v)
        fr = vals["copy_read_bytes"] / (vals["copy_fetch_kib"] * 1024)
        fw = vals["copy_write_bytes"] / (vals["copy_write_kib"] * 1024)
        return int(vals["fetch_kib"] * 1024 * fr + vals["write_kib"] * 1024 * fw)
    except Exception:
        return None


def timed(dev, steps, warmup, keep_x=False, **kw):
    for _ in range(warmup):
        dev.solve(**kw)
    runs = [dev.solve(keep_x=keep_x, **kw) for _ in range(steps)]
    d = runs[-1]
    d["t_solve_ms_median"] = statistics.median(r["t_solve_ms"] for r in runs)
    return d


def host_info():
    """The host the CPU baseline runs on: CPU model, sockets x cores (lscpu), the CPUs this
    process may use, and the OpenMP settings the reference runs with."""
    info = {"cpu_model": None, "physical_cores": None, "logical_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)),
            "omp": {k: os.environ.get(k) for k in ("OMP_NUM_THREADS", "OMP_PROC_BIND", "OMP_PLACES")}}
    try:
        import subprocess
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        kv = {ln.split(":", 1)[0].strip(): ln.split(":", 1)[1].strip() for ln in out.splitlines() if ":" in ln}
        info["cpu_model"] = kv.get("Model name")
        if kv.get("Socket(s)", "").isdigit() and kv.get("Core(s) per socket", "").isdigit():
            info["physical_cores"] = int(kv["Socket(s)"]) * int(kv["Core(s) per socket"])
    except Exception as e:  # pragma: no cover
        info["lscpu_error"] = str(e)
    return info


def ref_call(ref, fn):
    """Run a reference call with the process's stdout captured (the reference prints its
    iteration count and residual); returns (seconds, iterations, '# res' line)."""
    import tempfile
    sys.stdout.flush()
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            t0 = time.perf_counter()
            fn()
            dt = time.perf_counter() - t0
            ref.ref_flush()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        text = tmp.read().decode(errors="replace")
    iters, res = None, None
    for ln in text.splitlines():
        if ln.startswith("I iter ") or ln.startswith("# iter "):
            iters = int(ln.split()[2])
        elif ln.startswith("# res"):
            res = ln.strip()
    return dt, iters, res


def relres_of(sy, x):
    rows = np.repeat(np.arange(sy["N"]), np.diff(sy["rptr"]))
    r = sy["b"] - np.bincount(rows, weights=sy["val"] * x[sy["cind"]], minlength=sy["N"])
    return float(np.linalg.norm(r) / np.linalg.norm(sy["b"]))


def ref_abmc_gmres(dev, m, eps, budget_s, expect_iters, bcsr=None):
    """The reference's AbmcIluGmresm<double> (or, with bcsr=(bnl, bnw), AbmcIlubGmresm
    <double,bnl,bnw> on the BCSR copies of the factors) on the host cores, with the device's
    factors (equal to the reference's Ilu0 of the same pattern bit for bit).  One restart
    cycle runs first; the full solve runs when that rate, for 1.5x the expected iterations,
    fits budget_s seconds."""
    from oracle import oracle as O
    ref = O.ref()
    if ref is None:
        return {"error": "oracle/_ref not built"}
    sy, N = dev.sy, dev.N
    (lv, lc, lr), (uv, uc, ur) = dev.factors()
    if bcsr:
        bnl, bnw = bcsr
        L = S.csr2bcsr(lv, lc, lr, N, bnl, bnw)
        U = S.csr2bcsr(uv, uc, ur, N, bnl, bnw)
    else:
        L, U = (lv, lc, lr), (uv, uc, ur)
    b = sy["b"]
    nrm_b = float(np.linalg.norm(b))

    def run(outer, x):
        if bcsr:
            return lambda: ref.ref_abmc_ilub_gmresm_d(
                O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), *(O.P(t) for t in L), *(O.P(t) for t in U),
                O.P(sy["cptr"]), sy["nc"], sy["bsize"], O.P(b), O.P(x), nrm_b, outer, m, N, eps, bnl, bnw)
        return lambda: ref.ref_abmc_ilu_gmresm_d(
            O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), *(O.P(t) for t in L), *(O.P(t) for t in U),
            O.P(sy["cptr"]), sy["nc"], sy["bsize"], O.P(b), O.P(x), nrm_b, outer, m, N, eps)

    x = np.zeros(N)
    t_cycle, _, _ = ref_call(ref, run(1, x))
    out = {"ms_per_iteration_first_cycle": t_cycle * 1e3 / m}
    x = np.zeros(N)
    outer = 20000 // m
    fits = t_cycle / m * 1.5 * expect_iters <= budget_s
    t, iters, res = ref_call(ref, run(outer, x)) if fits else (None, None, None)
    if t is None:
        # the full solve does not fit the budget: one more cycle is the sample
        out.update({"full": False})
        return out
    out.update({"full": True, "t_solve_ms": t * 1e3, "iters": iters, "res_line": res,
                "ms_per_iteration": t * 1e3 / max(iters or 1, 1), "true_relres": relres_of(sy, x), "x": x})
    return out


def cpu_baseline(dev, gpu_iters, info):
    """The reference's AbmcIluGmresm (oracle/_ref, OpenMP) on the host: the full FP64
    ILU(0)-GMRES(30) solve of the same reordered system with the same factors (the device
    factors equal the reference's Ilu0 bit for bit), with the count and '# res' line it
    prints, plus 20 reference SpMVs."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    ref = O.ref()
    if ref is None:
        return {"error": "oracle/_ref not built"}
    sy, N = dev.sy, dev.N
    r = ref_abmc_gmres(dev, 30, 1e-10, 60.0, gpu_iters)
    y = np.zeros(N)
    xx = np.random.default_rng(0).uniform(-1, 1, N)
    ref.ref_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xx), O.P(y), N)
    t0 = time.perf_counter()
    for _ in range(20):
        ref.ref_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xx), O.P(y), N)
    t_spmv = (time.perf_counter() - t0) / 20
    threads = int(os.environ.get("OMP_NUM_THREADS") or info["affinity_cpus"])
    out = {"unit": "ms", "cores": threads, "kind": "reference", "cpu_model": info["cpu_model"],
           "physical_cores": info["physical_cores"], "omp": info["omp"],
           "spmv_GBps": spmv_bytes(N, sy["nnz"], 8) / t_spmv / 1e9}
    if r.get("full"):
        out.update({"value": r["t_solve_ms"], "ref_iters": r["iters"], "gpu_iters": gpu_iters,
                    "ref_res_line": r["res_line"], "ms_per_iteration": r["ms_per_iteration"],
                    "sample": (f"the full reference AbmcIluGmresm<double>(m=30) solve to 1e-10 of the same "
                               f"{round(N ** (1 / 3))}^3 system and ILU(0) factors on {threads} OpenMP threads "
                               f"(true relres {r['true_relres']:.2e})")})
    else:
        msi = r["ms_per_iteration_first_cycle"]
        out.update({"value": msi * gpu_iters, "ms_per_iteration": msi, "gpu_iters": gpu_iters,
                    "sample": (f"one restart cycle (30 iterations) of the reference AbmcIluGmresm<double>(m=30); "
                               f"value = per-iteration time x the {gpu_iters} GPU iterations")})
    return out


def load_learned(ctx, nested=False):
    """The trained correction: of the ILU application (flexible solvers), or with
    nested=True of the nested solve's inner result."""
    try:
        from senk_amd import gnn as G
        path = G.NESTED_WEIGHTS if nested else G.DEFAULT_WEIGHTS
        if not path.exists():
            return None, None
        theta, meta = G.load(path)
        return G.Gnn(ctx, theta), meta
    except Exception as e:  # pragma: no cover
        log(f"[bench] learned correction unavailable: {e}")
        return None, None


MIXED = dict(m=30, eps=1e-10, max_iter=6000, mixed=2, inner_eps=1e-3, outer_m=20)
# inner BiCGStab steps per outer step of the nested solve (sweeps over 40..120 steps,
# tolerances 3e-3..1e-4 and block sizes 4 / 8 / 16 on this system, with the split inner
# form and without: this setting is within noise of the best of either)
INNER_BICGSTAB = 60
# inner CG steps per outer step of the nested solve (split form; sweeps over 30..100 steps,
# tolerances 1e-2 / 1e-3 and block sizes 4 / 8: 30.8-47 ms at 128^3, this within noise of
# the best at blocks of 8)
INNER_CG = 100


def ingest_rates(n):
    """MatrixMarket ingestion on the host: the n^3 variable-coefficient matrix written as a
    symmetric coordinate file (its lower triangle), read by the library
    (senk_read_matrix_market + senk_expand) and by the reference's io::ReadMatrixMarket +
    Expand (oracle/_ref), median of 3 reads each; both must give the same arrays."""
    import tempfile
    rptr, cind, val = S.gen_vardiff7(n, 1, 8.0, 1.5)
    N = len(rptr) - 1
    rows = np.repeat(np.arange(N), np.diff(rptr))
    keep = cind <= rows
    r1, c1, v1 = rows[keep] + 1, cind[keep] + 1, val[keep]
    out = {"matrix": f"vardiff7 {n}^3 lower triangle", "N": N, "entries": int(keep.sum())}
    with tempfile.TemporaryDirectory() as d:
        path = os.path.join(d, "a.mtx")
        with open(path, "w") as f:
            f.write("%%MatrixMarket matrix coordinate real symmetric\n")
            f.write(f"{N} {N} {int(keep.sum())}\n")
            f.write("\n".join(f"{i} {j} {x!r}" for i, j, x in zip(r1.tolist(), c1.tolist(), v1.tolist())))
            f.write("\n")
        out["file_MB"] = round(os.path.getsize(path) / 1e6, 1)
        ts = []
        for _ in range(3):
            t0 = time.perf_counter()
            v, c, r, N_, _, shape = S.read_matrix_market(path)
            ev, ec, er = S.expand(v, c, r, N_, "L")
            ts.append(time.perf_counter() - t0)
        out["library_s"] = round(statistics.median(ts), 3)
        out["library_MBps"] = round(out["file_MB"] / out["library_s"], 1)
        try:
            from oracle import oracle as O
            ref = O.ref()
        except Exception:
            ref = None
        if ref is not None:
            cap = len(v) + 16
            rv, rc, rr, dims = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32), np.zeros(3, np.int32)
            xv, xc, xr = np.zeros(2 * cap), np.zeros(2 * cap, np.int32), np.zeros(N + 1, np.int32)
            ts = []
            for _ in range(3):
                t0 = time.perf_counter()
                k = ref.ref_read_mm(path.encode(), 1, O.P(rv), O.P(rc), O.P(rr), O.P(dims), cap, N + 1)
                ke = ref.ref_expand_d(O.P(rv), O.P(rc), O.P(rr), N, b"L", O.P(xv), O.P(xc), O.P(xr), 2 * cap)
                ts.append(time.perf_counter() - t0)
            out["reference_s"] = round(statistics.median(ts), 3)
            out["reference_MBps"] = round(out["file_MB"] / out["reference_s"], 1)
            out["identical"] = bool(k == len(v) and np.array_equal(rv[:k].view(np.int64), v.view(np.int64))
                                    and np.array_equal(xc[:ke], ec) and np.array_equal(xv[:ke], ev)
                                    and np.array_equal(xr, er))
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # median of 5: about one solve in four on the boxes measured here runs 5-8 ms long
    # (the same for every solver and host lag; not a property of the solve)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--ordering", choices=("connect", "simple"), default="connect",
                    help="--kernels-only / --learned-only: ABMC blocking of the blocks-of-8 system")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-c2", action="store_true")
    ap.add_argument("--no-c3", action="store_true", help="skip the unstructured 4,194,304-row workload")
    ap.add_argument("--no-c5", action="store_true", help="skip the 96^3 block (3-dof elasticity) workload")
    ap.add_argument("--no-ingest", action="store_true", help="skip the MatrixMarket ingestion entry")
    ap.add_argument("--kernels-only", action="store_true", help="kernel rates and the copy calibration only")
    ap.add_argument("--cold", action="store_true",
                    help="--kernels-only: also the single-launch rates after a 1 GiB read (the full run always)")
    ap.add_argument("--learned-only", action="store_true",
                    help="time the learned correction's application alone (FP32, 10 applications)")
    ap.add_argument("--run", default=None, help="time only the solve configuration of this name (profiling)")
    ap.add_argument("--sweep-modes", action="store_true",
                    help="also time the ILU application with one cooperative launch per triangle")
    ap.add_argument("--ingest", action="store_true",
                    help="host only: MatrixMarket reading, the library's reader against the reference's")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not supported (use --gpus 1)"}))
        sys.exit(2)
    if args.ingest:
        print(json.dumps(ingest_rates(args.n)), flush=True)
        return
    # the reference's OpenMP threads stay on their cores (read when libgomp starts, i.e.
    # when the compiled reference is first loaded)
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    n = args.n
    ctx = S.Context(0)
    log(f"[bench] building the {n}^3 variable-coefficient diffusion system (ABMC blocks of 8 and 4)")
    sy8 = build_system_kind("vardiff", n, 8, method=args.ordering if args.kernels_only or args.learned_only
                            else "connect")
    d8 = Device(ctx, sy8)
    if args.learned_only:
        # one layer pass per application: its kernels at this size for a trace or counters
        net, _ = load_learned(ctx)
        if net is None:
            raise SystemExit("no trained network")
        net.bind(d8.A)
        rng = np.random.default_rng(0)
        dr = ctx.array(rng.uniform(-1, 1, d8.N).astype(np.float32))
        dz = ctx.zeros(d8.N, np.float32)
        S.call("senk_prec_apply_f32", d8.M, dr.p, dz.p)
        us = 1e3 * time_kernel(ctx, lambda: S.call("senk_gnn_apply_f32", net.h, dr.p, dz.p), 10)
        print(json.dumps({"learned_apply_f32_us": us, "n": d8.N}), flush=True)
        net.close()
        d8.close()
        return
    if not args.run:  # (a --run trace holds that solve's kernels only)
        kern = kernel_rates(ctx, d8, sweep_modes=args.sweep_modes, cold=args.cold or not args.kernels_only)
        cal = copy_rate(ctx)
        log(f"[bench] kernels: {json.dumps(kern)} copy: {json.dumps(cal)}")
        if args.kernels_only:
            print(json.dumps({"kernels": kern, "copy": cal}), flush=True)
            return
    sy4 = build_system_kind("vardiff", n, 4)
    d4 = Device(ctx, sy4)
    # "simple" ABMC blocking (blocks of 8 consecutive rows): the block graph of a 7-point
    # grid is 2-colorable, so each triangular solve has 2 levels instead of 10
    sys8 = build_system_kind("vardiff", n, 8, method="simple")
    ds8 = Device(ctx, sys8)
    K, W = args.steps, args.warmup
    if args.run:
        # one configuration alone (for a kernel trace of exactly that solve)
        cfgs = {"fp64_gmres30_b8": (d8, dict(solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000)),
                "mixed_bicgstab_b4": (d4, dict(solver=S.SENK_BICGSTAB, inner_max=INNER_BICGSTAB, **MIXED)),
                "mixed_bicgstab_b8": (d8, dict(solver=S.SENK_BICGSTAB, inner_max=INNER_BICGSTAB, **MIXED)),
                "mixed_cg_b8": (d8, dict(solver=S.SENK_CG, inner_max=INNER_CG, **MIXED)),
                "mixed_cg_s8": (ds8, dict(solver=S.SENK_CG, inner_max=INNER_CG, **MIXED)),
                "mixed_bicgstab_s8": (ds8, dict(solver=S.SENK_BICGSTAB, inner_max=INNER_BICGSTAB, **MIXED)),
                "learned_fgcr30_fp64_b8": (d8, dict(solver=S.SENK_FGCR, m=30, eps=1e-10, max_iter=6000,
                                                    check_every=1, use_learned=1)),
                "learned_mixed_fgcr_b8": (d8, dict(solver=S.SENK_FGCR, inner_max=30, use_learned=1, **MIXED)),
                "learned_mixed_cg_b8": (d8, dict(solver=S.SENK_CG, inner_max=INNER_CG, use_learned=1, **MIXED)),
                "learned_mixed_cg_s8": (ds8, dict(solver=S.SENK_CG, inner_max=INNER_CG, use_learned=1, **MIXED))}
        if args.run == "fp64_ic0_cg_b8":
            d4.close()
            d4 = Device(ctx, d8.sy, kind=S.SENK_IC0)  # the IC(0) factor of the same system
            cfgs[args.run] = (d4, dict(solver=S.SENK_CG, eps=1e-10, max_iter=6000))
        dev, kw = cfgs[args.run]
        net = None
        if kw.get("use_learned"):
            net, _ = load_learned(ctx, nested=args.run.startswith("learned_mixed_cg"))
            if net is None:
                raise SystemExit("no trained network for " + args.run)
            net.bind(dev.A)
            net.attach(dev.M)
        r = timed(dev, K, W, **kw)
        if net is not None:
            S.call("senk_prec_attach_gnn", dev.M, None)
            net.close()
        print(json.dumps({"run": args.run, **r}), flush=True)
        d8.close()
        d4.close()
        ds8.close()
        return
    runs = {}
    # plain FP64 (the reference's method) at both block sizes; the baseline is the faster
    for tag, dev in (("b8", d8), ("b4", d4)):
        runs[f"fp64_gmres30_{tag}"] = timed(dev, K, W, solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000)
        log(f"[bench] fp64_gmres30_{tag}: {json.dumps(slim(runs[f'fp64_gmres30_{tag}']))}")
    # the same GMRES(30) with two-pass classical Gram-Schmidt (not the reference's MGS)
    runs["fp64_gmres30_cgs2_b8"] = timed(d8, K, W, solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000,
                                         ortho=S.SENK_CGS2)
    runs["fp64_bicgstab_b8"] = timed(d8, K, W, solver=S.SENK_BICGSTAB, m=30, eps=1e-10, max_iter=6000)
    # conjugate gradients with IC(0) (the reference announces both, has neither)
    for tag, sy_ in (("b8", d8.sy), ("s8", ds8.sy)):
        dic = Device(ctx, sy_, kind=S.SENK_IC0)
        runs[f"fp64_ic0_cg_{tag}"] = timed(dic, K, W, solver=S.SENK_CG, eps=1e-10, max_iter=6000)
        log(f"[bench] fp64_ic0_cg_{tag}: {json.dumps(slim(runs[f'fp64_ic0_cg_{tag}']))}")
        dic.close()
    runs["fp64_bicgstab_s8"] = timed(ds8, K, W, solver=S.SENK_BICGSTAB, m=30, eps=1e-10, max_iter=6000)
    runs["fp64_gmres30_s8"] = timed(ds8, K, W, solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000)
    for tag, dev in (("b8", d8), ("b4", d4), ("s8", ds8)):
        runs[f"mixed_bicgstab_{tag}"] = timed(dev, K, W, solver=S.SENK_BICGSTAB, inner_max=INNER_BICGSTAB, **MIXED)
        runs[f"mixed_fgcr_{tag}"] = timed(dev, K, W, solver=S.SENK_FGCR, inner_max=30, **MIXED)
        runs[f"mixed_cg_{tag}"] = timed(dev, K, W, solver=S.SENK_CG, inner_max=INNER_CG, **MIXED)
        log(f"[bench] mixed_{tag}: {json.dumps(slim(runs[f'mixed_cg_{tag}']))}")
    net, meta = load_learned(ctx)
    if net is not None:
        net.bind(d8.A)
        net.attach(d8.M)
        runs["learned_fgcr30_fp64_b8"] = timed(d8, K, W, solver=S.SENK_FGCR, m=30, eps=1e-10, max_iter=6000,
                                               check_every=1, use_learned=1)
        runs["plain_fgcr30_fp64_b8"] = timed(d8, K, W, solver=S.SENK_FGCR, m=30, eps=1e-10, max_iter=6000,
                                             check_every=1)
        runs["learned_mixed_fgcr_b8"] = timed(d8, K, W, solver=S.SENK_FGCR, inner_max=30, use_learned=1, **MIXED)
        log(f"[bench] learned_mixed_fgcr_b8: {json.dumps(slim(runs['learned_mixed_fgcr_b8']))}")
        dx32, dz32 = ctx.zeros(d8.N, np.float32), ctx.zeros(d8.N, np.float32)
        kern["learned_apply_f32_us"] = 1e3 * time_kernel(
            ctx, lambda: S.call("senk_gnn_apply_f32", net.h, dx32.p, dz32.p), 10)
        S.call("senk_prec_attach_gnn", d8.M, None)
        net.close()
    # the correction of the nested solve's inner result (FP64 flexible GCR around the FP32
    # split inner CG), trained on that solve's own (residual, inner result) pairs
    netn, meta_n = load_learned(ctx, nested=True)
    if netn is not None:
        for tag, dev in (("b8", d8), ("s8", ds8)):
            netn.bind(dev.A)
            netn.attach(dev.M)
            runs[f"learned_mixed_cg_{tag}"] = timed(dev, K, W, solver=S.SENK_CG, inner_max=INNER_CG, use_learned=1,
                                                    **MIXED)
            log(f"[bench] learned_mixed_cg_{tag}: {json.dumps(slim(runs[f'learned_mixed_cg_{tag}']))}")
            S.call("senk_prec_attach_gnn", dev.M, None)
        netn.close()

    def ok(r):  # a solve that counts: converged, FP64 true residual at the target
        return r["converged"] == 1 and r["true_relres"] <= 1e-10

    # the FP64 baseline: the fastest of the reference's method over the orderings tried
    base_key = min((k for k in ("fp64_gmres30_b8", "fp64_gmres30_b4", "fp64_gmres30_s8") if ok(runs[k])),
                   key=lambda k: runs[k]["t_solve_ms_median"])
    mixed_key = min((k for k in runs if k.startswith(("mixed_", "learned_mixed_")) and ok(runs[k])),
                    key=lambda k: runs[k]["t_solve_ms_median"])
    base, best = runs[base_key], runs[mixed_key]
    # the FP64 solve of the same method family as the headline solve (CG with IC(0) for the
    # inner-CG forms, BiCGStab for the inner-BiCGStab forms, FGCR for the FGCR forms)
    family = {"cg": ("fp64_ic0_cg_b8", "fp64_ic0_cg_s8"), "bicgstab": ("fp64_bicgstab_b8", "fp64_bicgstab_s8"),
              "fgcr": ("plain_fgcr30_fp64_b8",)}
    fam = next((v for k, v in family.items() if f"_{k}_" in mixed_key + "_"), ())
    fam_key = min((k for k in fam if k in runs and ok(runs[k])), key=lambda k: runs[k]["t_solve_ms_median"],
                  default=None)
    info = host_info()
    cpu = None
    devs = {"b8": d8, "b4": d4, "s8": ds8}
    if not args.no_cpu:
        log("[bench] CPU baseline (reference OpenMP AbmcIluGmresm)")
        cpu = cpu_baseline(devs[base_key[-2:]], base["iters"], info)
        log(f"[bench] cpu: {json.dumps(cpu)}")
    # the roofline entries describe the kernels of the headline solve's system
    head = devs[mixed_key[-2:]]
    hkern = kern if head is d8 else kernel_rates(ctx, head)
    hname = {"b8": "ABMC connect, blocks of 8", "b4": "ABMC connect, blocks of 4",
             "s8": "ABMC simple, blocks of 8"}[mixed_key[-2:]]
    hprof = {"b8": "vardiff128", "b4": "vardiff128_b4", "s8": "vardiff128_simple8"}[mixed_key[-2:]]
    d8.close()
    d4.close()
    ds8.close()
    Ks = min(K, 5)  # the other workloads: median of at most 5 solves each
    c2 = None if args.no_c2 else config_c2(ctx, n, Ks, args.no_cpu)
    c3 = None if args.no_c3 else config_c3(ctx, Ks, args.no_cpu)
    c5 = None if args.no_c5 else config_c5(ctx, Ks, args.no_cpu)
    value = best["t_solve_ms_median"]
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": K, "warmup": W,
        "ms_per_step": round(value, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner solve + f64 outer iteration", "data": "synthetic",
        "config": {"workload": f"vardiff7 {n}^3 (k = exp(GRF, corr 8, sigma 1.5), seed 1), ILU(0), tol 1e-10",
                   "n": sy8["N"], "nnz": sy8["nnz"], "mixed": mixed_key, "mixed_ordering": hname,
                   "fp64_baseline": base_key},
        "speedup_vs_fp64_gmres30": round(base["t_solve_ms_median"] / value, 3),
        "speedup_vs_fp64_same_method": None if fam_key not in runs else {
            "fp64": fam_key, "speedup": round(runs[fam_key]["t_solve_ms_median"] / value, 3)},
        "learned_vs_fp64_gmres30": {k: round(base["t_solve_ms_median"] / runs[k]["t_solve_ms_median"], 3)
                                    for k in ("learned_mixed_fgcr_b8", "learned_mixed_cg_b8", "learned_mixed_cg_s8")
                                    if k in runs},
        "runs": {k: slim(v) for k, v in runs.items()},
        "kernels": kern, "kernels_headline_system": hkern, "copy_calibration": cal,
        "learned_training": {name: {k: m_[k] for k in ("mode", "count", "sizes", "epochs", "loss_hist", "t_train_s",
                                                       "t_data_s", "n_residuals") if k in m_}
                             for name, m_ in (("ilu", meta), ("nested", meta_n)) if m_ is not None},
        "plain_fp64_c2": c2, "mixed_c3": c3, "block_c5": c5,
        "setup_s": {"abmc": sy8["t_abmc"], "reorder": sy8["t_reorder"], "upload": d8.t_upload,
                    "device_ilu0_split_layout": d8.t_ilu},
        # the dominant kernels of the headline solve: the FP32 ILU(0) application (one L sweep
        # and one U sweep; with the split inner form they are most of an inner step), and the
        # FP32 SELL SpMV beside it
        "roofline": {"kernel": f"ilu0_apply_f32: L + U sweeps, per-XCD ticket streams (vardiff7 {n}^3, {hname})",
                     "bound": "hbm", "achieved": round(hkern["ilu_apply_f32_GBps"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(hkern["ilu_apply_f32_GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(f"ilu_apply_f32_{hprof}.md"),
                     "achieved_cold": round(hkern["cold"]["ilu_apply_f32_GBps"], 1),
                     "frac_cold": round(hkern["cold"]["ilu_apply_f32_GBps"] / HBM_PEAK_GBS, 4)},
        "roofline_spmv": {"kernel": f"spmv_sellcs_f32 (vardiff7 {n}^3, {hname})", "bound": "hbm",
                          "achieved": round(hkern["spmv_sell_f32_GBps"], 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                          "frac": round(hkern["spmv_sell_f32_GBps"] / HBM_PEAK_GBS, 4),
                          "traffic": traffic_from_profile(f"spmv_sell_f32_{hprof}.md"),
                          "achieved_cold": round(hkern["cold"]["spmv_sell_f32_GBps"], 1),
                          "frac_cold": round(hkern["cold"]["spmv_sell_f32_GBps"] / HBM_PEAK_GBS, 4)},
        "cpu_baseline": cpu,
        "host": info,
        # host-side MatrixMarket ingestion (64^3 file; --ingest --n 128 for the full size)
        "ingest": None if args.no_ingest else ingest_rates(64),
    }
    print(json.dumps(line), flush=True)


KEEP = ("iters", "restarts", "inner_iters", "converged", "fell_back", "t_solve_ms_median", "true_relres",
        "err_vs_xstar", "err_vs_fp64_solve")


def slim(d):
    return {q: d[q] for q in KEEP if q in d}


def relerr(x, y):
    return float(np.linalg.norm(x - y) / np.linalg.norm(y))


def within5(gpu, ref):
    """The plain path's count criterion: within 5% of the reference (at least +-2)."""
    return None if ref is None else bool(abs(gpu - ref) <= max(2, int(np.ceil(0.05 * ref))))


def reference_entry(r, gpu_iters):
    if "error" in r:
        return r
    out = {k: r[k] for k in ("full", "t_solve_ms", "iters", "res_line", "ms_per_iteration",
                             "ms_per_iteration_first_cycle", "true_relres") if k in r}
    out["gpu_iters"] = gpu_iters
    out["iters_within_5pct"] = within5(gpu_iters, r.get("iters"))
    return out


# the nested mixed BiCGStab on the 27-point system: 120 inner steps to 1e-4 (sweeps of
# 60 / 120 / 300 steps to 1e-3 / 1e-4 / 1e-6, and GMRES(30) inner solves: this the fastest)
C2_INNER = dict(inner_max=120, inner_eps=1e-4)


def config_c2(ctx, n, K, no_cpu):
    """27-point convection-diffusion 128^3 (Pe 10), ABMC blocks of 16: plain FP64
    ILU(0)-GMRES(30) and ILU(0)-BiCGStab, the nested mixed form, the FP64 BiCGStab with the
    FP32 factor application, SpMV and ILU rates, and the
    reference's AbmcIluGmresm count on the same system and factors; then the same solves
    with "simple" blocks of 16 (8 colors instead of 13)."""
    log(f"[bench] building the {n}^3 27-point convection-diffusion system (ABMC blocks of 16)")
    sy = build_system_kind("convdiff", n, 16)
    dc = Device(ctx, sy)
    out = {"workload": f"convdiff27 {n}^3 (Pe 10), ABMC connect blocks of 16, ILU(0), tol 1e-10",
           "n": sy["N"], "nnz": sy["nnz"], "kernels": kernel_rates(ctx, dc)}
    mixed = dict(MIXED, **C2_INNER)
    out["fp64_gmres30"] = slim(timed(dc, K, 1, solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000))
    out["fp64_bicgstab"] = slim(timed(dc, K, 1, solver=S.SENK_BICGSTAB, eps=1e-10, max_iter=6000))
    out["mixed_bicgstab"] = slim(timed(dc, K, 1, solver=S.SENK_BICGSTAB, **mixed))
    # the FP64 BiCGStab with its ILU(0) application on the FP32 factor copies (mixed = 3)
    out["fp64_bicgstab_f32_ilu"] = slim(timed(dc, K, 1, solver=S.SENK_BICGSTAB, eps=1e-10, max_iter=6000, mixed=3))
    if not no_cpu:
        gi = out["fp64_gmres30"]["iters"]
        out["reference_gmres30"] = reference_entry(ref_abmc_gmres(dc, 30, 1e-10, 40.0, gi), gi)
    dc.close()
    ss = build_system_kind("convdiff", n, 16, method="simple")
    ds = Device(ctx, ss)
    out["simple16"] = {"colors": ss["nc"], "kernels": kernel_rates(ctx, ds, cold=False),
                       "fp64_gmres30": slim(timed(ds, K, 1, solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000)),
                       "fp64_bicgstab": slim(timed(ds, K, 1, solver=S.SENK_BICGSTAB, eps=1e-10, max_iter=6000)),
                       "mixed_bicgstab": slim(timed(ds, K, 1, solver=S.SENK_BICGSTAB, **mixed))}
    ds.close()
    log(f"[bench] c2: {json.dumps(out)}")
    return out


def config_c3(ctx, K, no_cpu, n=4194304):
    """Unstructured nonsymmetric matrix, n = 4,194,304: FP32 values and ILU(0) factors,
    FP32 SpMV and inner GMRES(50), FP64 refinement to 1e-10; error against an FP64 solve
    of the same system (the device's FP64 ILU(0)-GMRES(50) to 1e-14); the plain FP64
    GMRES(50), the FP64 FGMRES(50) with the FP32 factor application, and the reference's
    AbmcIluGmresm(m = 50) count beside it."""
    log(f"[bench] building the unstructured system (n = {n}, ABMC blocks of 16)")
    sy = build_system_kind("unstructured", n, 16)
    dev = Device(ctx, sy)
    out = {"workload": f"unstructured n={n} (lognormal rows, banded-random columns, diagonal 1.1x row sum), ABMC "
                       f"connect blocks of 16 ({sy['nc']} colors), ILU(0), tol 1e-10",
           "n": sy["N"], "nnz": sy["nnz"], "setup_s": {"generate": sy["t_gen"], "abmc": sy["t_abmc"],
                                                      "reorder": sy["t_reorder"], "device": dev.t_ilu},
           "kernels": kernel_rates(ctx, dev)}
    r64 = dev.solve(keep_x=True, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
    x64 = r64.pop("x")
    out["fp64_solve_1e-14"] = slim(r64)
    out["fp64_gmres50"] = slim(timed(dev, K, 1, solver=S.SENK_GMRES, m=50, eps=1e-10, max_iter=20000))
    m = timed(dev, K, 1, keep_x=True, solver=S.SENK_GMRES, m=50, eps=1e-10, max_iter=20000, mixed=1,
              inner_eps=1e-5, inner_max=500)
    m["err_vs_fp64_solve"] = relerr(m.pop("x"), x64)
    out["mixed_ir_gmres50"] = slim(m)
    # FP64 FGMRES(50) with the ILU(0) application on the FP32 factor copies (mixed = 3)
    m = timed(dev, K, 1, keep_x=True, solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000, mixed=3)
    m["err_vs_fp64_solve"] = relerr(m.pop("x"), x64)
    out["fp64_fgmres50_f32_ilu"] = slim(m)
    if not no_cpu:
        gi = out["fp64_gmres50"]["iters"]
        out["reference_gmres50"] = reference_entry(ref_abmc_gmres(dev, 50, 1e-10, 40.0, gi), gi)
    log(f"[bench] c3: {json.dumps(out)}")
    dev.close()
    return out


def config_c5(ctx, K, no_cpu, n=96):
    """Block path: 3 dof per node on n^3 (3x3-block 7-point stencil, nu 0.3, lognormal
    moduli), ABMC blocks of 48: ILUB (3x3 block fill) and ILU(1) over the colors, FP32
    factors + FP64 refinement + FGMRES(50), with ILUB also with the learned correction; error
    against the FP64 ILUB-GMRES(50) solve to 1e-14; the FP64 FGMRES(50) with the factor
    application on the FP32 copies; the plain FP64 ILUB-GMRES(50) and the
    reference's AbmcIlubGmresm<double,3,3> count beside it."""
    log(f"[bench] building the {n}^3 3-dof block system (ABMC blocks of 48)")
    sy = build_system_kind("elasticity", n, 48)
    out = {"workload": f"elasticity3 {n}^3 (3 dof, nu 0.3, lognormal E), ABMC connect blocks of 48 "
                       f"({sy['nc']} colors), tol 1e-10", "n": sy["N"], "nnz": sy["nnz"]}
    mixed = dict(solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000, mixed=1, inner_eps=1e-5, inner_max=500)
    x64 = None
    for tag, kw in (("ilub", dict(kind=S.SENK_ILUB, bnl=3, bnw=3)), ("ilu1", dict(kind=S.SENK_ILUP, p=1))):
        dev = Device(ctx, sy, **kw)
        out[f"setup_{tag}_s"] = dev.t_ilu
        if x64 is None:
            r64 = dev.solve(keep_x=True, solver=S.SENK_GMRES, m=50, eps=1e-14, max_iter=20000)
            x64 = r64.pop("x")
            out["fp64_ilub_solve_1e-14"] = slim(r64)
            out["fp64_ilub_gmres50"] = slim(timed(dev, K, 1, solver=S.SENK_GMRES, m=50, eps=1e-10, max_iter=20000))
            if not no_cpu:
                gi = out["fp64_ilub_gmres50"]["iters"]
                out["reference_ilub_gmres50"] = reference_entry(
                    ref_abmc_gmres(dev, 50, 1e-10, 40.0, gi, bcsr=(3, 3)), gi)
        r = timed(dev, K, 1, keep_x=True, **mixed)
        r["err_vs_fp64_solve"] = relerr(r.pop("x"), x64)
        out[f"mixed_{tag}_fgmres50"] = slim(r)
        # FP64 FGMRES(50) with the factor application on the FP32 copies (mixed = 3)
        r = timed(dev, K, 1, keep_x=True, solver=S.SENK_FGMRES, m=50, eps=1e-10, max_iter=20000, mixed=3)
        r["err_vs_fp64_solve"] = relerr(r.pop("x"), x64)
        out[f"fp64_{tag}_fgmres50_f32_apply"] = slim(r)
        net, _ = load_learned(ctx) if tag == "ilub" else (None, None)
        if net is not None:
            net.bind(dev.A)
            net.attach(dev.M)
            r = timed(dev, min(K, 3), 1, keep_x=True, use_learned=1, **mixed)
            r["err_vs_fp64_solve"] = relerr(r.pop("x"), x64)
            out[f"learned_mixed_{tag}_fgmres50"] = slim(r)
            S.call("senk_prec_attach_gnn", dev.M, None)
            net.close()
        dev.close()
    log(f"[bench] c5: {json.dumps(out)}")
    return out


if __name__ == "__main__":
    main()
