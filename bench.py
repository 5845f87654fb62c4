"""Benchmark of SenK's preconditioned Krylov hot path on one MI355X.

Workload (one GPU): the 3D 27-point convection-diffusion system on a 128^3 grid
(n = 2,097,152, nnz = 55,742,968; unit diffusion, central convection with velocity
Pe*(1, 0.5, 0.25), Pe = 10), ABMC ordering with blocks of 16 ("connect"), ILU(0),
right-hand side b = A x* with x* ~ U[-1,1] from seed 1, x0 = 0, target 1e-10 relative
residual.  One "step" is one full solve with all inputs resident in HBM (setup is timed
separately, as the reference's driver times only the solve).

Reported on the one JSON line:
  value        time-to-1e-10 of the mixed-precision solve (FP32 matrix/ILU(0) factors and
               inner GMRES(30), FP64 iterative refinement), ms per solve
  fp64_gmres   the plain FP64 ILU(0)-GMRES(30) solve (time, iterations, true residual)
  fp64_bicgstab / mixed details / SpMV GB/s (CSR and SELL-C-sigma, FP64 and FP32)
  roofline     FP64 CSR SpMV: algorithmic bytes / average launch time (HIP events on the
               library's stream) against 8 TB/s
  cpu_baseline the reference's own OpenMP AbmcIluGmresm (oracle/_ref) on the host cores

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--n 128]
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


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, sv, si=4):
    return nnz * (sv + si) + (n + 1) * si + 2 * n * sv


def build_system(n, bsize):
    return build_system_kind("convdiff", n, bsize)


def build_system_kind(kind, n, bsize, seed=1):
    """kind "convdiff": 27-point convection-diffusion, Pe 10; "vardiff": 7-point
    -div(k grad u) with k = exp(Gaussian random field, correlation length 8, sigma 1.5)."""
    t0 = time.perf_counter()
    if kind == "convdiff":
        rptr, cind, val = S.gen_convdiff27(n, 10.0)
    elif kind == "vardiff":
        rptr, cind, val = S.gen_vardiff7(n, seed, 8.0, 1.5)
    else:
        raise ValueError(kind)
    N = len(rptr) - 1
    xs = S.gen_uniform(1, N)
    t1 = time.perf_counter()
    nc, cptr, LP, RP = S.abmc_permutation(cind, rptr, N, bsize, False, "connect")
    t2 = time.perf_counter()
    val2, cind2, rptr2 = S.reordering(val, cind, rptr, LP, RP, N)
    t3 = time.perf_counter()
    # b = A x* in the original ordering, then permuted with the rows
    b = _host_spmv(rptr, cind, val, xs)
    bp = b[LP].copy()
    return dict(N=N, nnz=int(rptr[-1]), rptr=rptr2, cind=cind2, val=val2, b=bp, xs=xs[LP].copy(), nc=nc,
                cptr=cptr, LP=LP, t_gen=t1 - t0, t_abmc=t2 - t1, t_reorder=t3 - t2)


def _host_spmv(rptr, cind, val, x):
    rows = np.repeat(np.arange(len(rptr) - 1), np.diff(rptr))
    return np.bincount(rows, weights=val * x[cind], minlength=len(rptr) - 1)


def time_kernel(ctx, fn, reps):
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


def solve(ctx, A, M, db, N, **kw):
    dx = ctx.zeros(N)
    o = S.solve_opts(**kw)
    res = S.senk_result()
    t0 = time.perf_counter()
    S.call("senk_solve", ctx.h, A, M, db.p, dx.p, C.byref(o), C.byref(res), None, 0)
    wall = (time.perf_counter() - t0) * 1e3
    d = res.as_dict()
    d["wall_ms"] = wall
    return d, dx


def cpu_baseline(sy, ctx, M, gpu_iters):
    """The reference's AbmcIluGmresm (oracle/_ref) on the host: one restart cycle of 30
    iterations on the same system and factors (the device factors equal the reference's
    Ilu0 bit for bit), scaled to the GPU run's iteration count; plus 20 reference SpMVs."""
    try:
        from oracle import oracle as O
    except Exception as e:  # pragma: no cover
        return {"error": f"oracle unavailable: {e}"}
    ref = O.ref()
    kind = "reference"
    if ref is None:
        return {"error": "oracle/_ref not built"}
    N = sy["N"]
    ln, un = C.c_int(), C.c_int()
    S.call("senk_prec_factor_nnz", M, C.byref(ln), C.byref(un), None, None)
    lr, lc, lv = np.zeros(N + 1, np.int32), np.zeros(ln.value, np.int32), np.zeros(ln.value)
    ur, uc, uv = np.zeros(N + 1, np.int32), np.zeros(un.value, np.int32), np.zeros(un.value)
    S.call("senk_prec_download", M, O.P(lr), O.P(lc), O.P(lv), O.P(ur), O.P(uc), O.P(uv))
    x = np.zeros(N)
    b = sy["b"]
    nrm_b = float(np.linalg.norm(b))
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    # silence the reference's stdout lines while it runs
    sys.stdout.flush()
    saved = os.dup(1)
    devnull = os.open(os.devnull, os.O_WRONLY)
    os.dup2(devnull, 1)
    try:
        t0 = time.perf_counter()
        ref.ref_abmc_ilu_gmresm_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(lv), O.P(lc), O.P(lr),
                                  O.P(uv), O.P(uc), O.P(ur), O.P(sy["cptr"]), sy["nc"], 16, O.P(b), O.P(x), nrm_b,
                                  1, 30, N, 1e-10)
        t_cycle = time.perf_counter() - t0
        y = np.zeros(N)
        xx = np.random.default_rng(0).uniform(-1, 1, N)
        ref.ref_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xx), O.P(y), N)
        t0 = time.perf_counter()
        reps = 20
        for _ in range(reps):
            ref.ref_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xx), O.P(y), N)
        t_spmv = (time.perf_counter() - t0) / reps
        ref.ref_flush()
    finally:
        os.dup2(saved, 1)
        os.close(saved)
        os.close(devnull)
    per_iter_ms = t_cycle * 1e3 / 30
    return {
        "value": per_iter_ms * gpu_iters, "unit": "ms", "cores": cores, "kind": kind,
        "sample": f"one restart cycle (30 iterations) of the reference AbmcIluGmresm<double>(m=30) on the same "
                  f"128^3 system and ILU(0) factors took {t_cycle:.2f} s; value = per-iteration time x the "
                  f"{gpu_iters} iterations the FP64 GPU solve needed",
        "ms_per_iteration": per_iter_ms,
        "spmv_GBps": spmv_bytes(N, sy["nnz"], 8) / t_spmv / 1e9,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=128)
    ap.add_argument("--bsize", type=int, default=16)
    ap.add_argument("--no-cpu", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU runs are not supported (use --gpus 1)"}))
        sys.exit(2)
    n, bsize = args.n, args.bsize
    log(f"[bench] building {n}^3 27-point convection-diffusion system")
    sy = build_system(n, bsize)
    N, nnz = sy["N"], sy["nnz"]
    log(f"[bench] N={N} nnz={nnz} colors={sy['nc']} (gen {sy['t_gen']:.1f}s abmc {sy['t_abmc']:.1f}s "
        f"reorder {sy['t_reorder']:.1f}s)")
    ctx = S.Context(0)
    A = C.c_void_p()
    t0 = time.perf_counter()
    S.call("senk_mat_create", ctx.h, N, S._ptr(sy["rptr"]), S._ptr(sy["cind"]), S._ptr(sy["val"]), 1 | 2, C.byref(A))
    t_upload = time.perf_counter() - t0
    M = C.c_void_p()
    opts = S.senk_ilu_opts(S.SENK_ILU0, 0, 0, 0, (1 << S.SENK_FACTOR_F64) | (1 << S.SENK_FACTOR_F32))
    t0 = time.perf_counter()
    S.call("senk_prec_create_ilu", ctx.h, A, S._ptr(sy["rptr"]), S._ptr(sy["cind"]), S._ptr(sy["val"]),
           S._ptr(sy["cptr"]), sy["nc"], bsize, C.byref(opts), C.byref(M))
    t_ilu = time.perf_counter() - t0
    log(f"[bench] upload {t_upload:.2f}s, device ILU(0) setup {t_ilu:.2f}s")
    db = ctx.array(sy["b"])

    # ---- kernel rates (HIP events on the library stream)
    dx = ctx.array(np.random.default_rng(0).uniform(-1, 1, N))
    dy = ctx.zeros(N)
    dr, dc, dv = ctx.array(sy["rptr"]), ctx.array(sy["cind"]), ctx.array(sy["val"])
    reps = 50
    t_csr = time_kernel(ctx, lambda: S.call("senk_spmv_csr_f64", ctx.h, dv.p, dc.p, dr.p, dx.p, dy.p, N), reps)
    t_mat = time_kernel(ctx, lambda: S.call("senk_mat_spmv_f64", A, dx.p, dy.p), reps)
    dx32, dy32 = ctx.zeros(N, np.float32), ctx.zeros(N, np.float32)
    t_mat32 = time_kernel(ctx, lambda: S.call("senk_mat_spmv_f32", A, dx32.p, dy32.p), reps)
    t_prec = time_kernel(ctx, lambda: S.call("senk_prec_apply_f64", M, dx.p, dy.p), 10)
    t_prec32 = time_kernel(ctx, lambda: S.call("senk_prec_apply_f32", M, dx32.p, dy32.p), 10)
    ln, un, nl, nu = C.c_int(), C.c_int(), C.c_int(), C.c_int()
    S.call("senk_prec_factor_nnz", M, C.byref(ln), C.byref(un), C.byref(nl), C.byref(nu))
    b64 = spmv_bytes(N, nnz, 8)
    b32 = spmv_bytes(N, nnz, 4)
    ilu_bytes64 = (ln.value + un.value) * 12 + 2 * (N + 1) * 4 + 4 * N * 8
    ilu_bytes32 = (ln.value + un.value) * 8 + 2 * (N + 1) * 4 + 4 * N * 4
    del dv, dc, dr
    kern = {
        "spmv_csr_f64_us": t_csr * 1e3, "spmv_csr_f64_GBps": b64 / (t_csr * 1e-3) / 1e9,
        "spmv_sellcs_f64_us": t_mat * 1e3, "spmv_sellcs_f64_GBps": b64 / (t_mat * 1e-3) / 1e9,
        "spmv_sellcs_f32_us": t_mat32 * 1e3, "spmv_sellcs_f32_GBps": b32 / (t_mat32 * 1e-3) / 1e9,
        "ilu_apply_f64_us": t_prec * 1e3, "ilu_apply_f64_GBps": ilu_bytes64 / (t_prec * 1e-3) / 1e9,
        "ilu_apply_f32_us": t_prec32 * 1e3, "ilu_apply_f32_GBps": ilu_bytes32 / (t_prec32 * 1e-3) / 1e9,
        "colors": sy["nc"],
    }
    log(f"[bench] kernels: {json.dumps(kern)}")

    # ---- solves
    runs = {"fp64_gmres": dict(solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000),
            "fp64_gmres_cgs2": dict(solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000, ortho=S.SENK_CGS2),
            "fp64_bicgstab": dict(solver=S.SENK_BICGSTAB, m=30, eps=1e-10, max_iter=6000),
            "mixed_gmres": dict(solver=S.SENK_GMRES, m=30, eps=1e-10, max_iter=6000, mixed=1, inner_eps=1e-5,
                                inner_max=600, ortho=S.SENK_CGS2)}
    results = {}
    for name, kw in runs.items():
        for _ in range(args.warmup):
            solve(ctx, A, M, db, N, **kw)
        times, last = [], None
        for _ in range(args.steps if name == "mixed_gmres" else max(1, min(args.steps, 2))):
            d, dxs = solve(ctx, A, M, db, N, **kw)
            times.append(d["t_solve_ms"])
            last = d
        x = dxs.download()
        last["err_vs_xstar"] = float(np.linalg.norm(x - sy["xs"]) / np.linalg.norm(sy["xs"]))
        last["t_solve_ms_median"] = statistics.median(times)
        results[name] = last
        log(f"[bench] {name}: {json.dumps(last)}")
    # error of the mixed solution against the FP64 solve
    ref64 = results["fp64_gmres"]
    mixed = results["mixed_gmres"]
    cpu = None
    if not args.no_cpu:
        log("[bench] CPU baseline (reference OpenMP path)")
        cpu = cpu_baseline(sy, ctx, M, ref64["iters"])
        log(f"[bench] cpu: {json.dumps(cpu)}")

    value = mixed["t_solve_ms_median"]
    line = {
        "metric": "time-to-1e-10 residual & iters: mixed/learned vs FP64 ILU(0)-GMRES; SpMV GB/s",
        "value": round(value, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(value, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32 inner + f64 refinement (fp64 path reported beside it)", "data": "synthetic",
        "config": {"workload": f"convdiff27 {n}^3 Pe=10, ABMC bsize {bsize} connect, ILU(0), tol 1e-10",
                   "n": N, "nnz": nnz},
        "fp64_gmres30": {k: ref64[k] for k in ("iters", "t_solve_ms_median", "true_relres", "err_vs_xstar")},
        "fp64_gmres30_cgs2": {k: results["fp64_gmres_cgs2"][k] for k in ("iters", "t_solve_ms_median", "true_relres")},
        "fp64_bicgstab": {k: results["fp64_bicgstab"][k] for k in ("iters", "t_solve_ms_median", "true_relres")},
        "mixed_ir_gmres30": {k: mixed[k] for k in ("iters", "restarts", "t_solve_ms_median", "true_relres",
                                                     "err_vs_xstar")},
        "speedup_mixed_vs_fp64_gmres": round(ref64["t_solve_ms_median"] / value, 3),
        "kernels": kern,
        "setup_s": {"abmc": sy["t_abmc"], "reorder": sy["t_reorder"], "device_ilu0_and_split": t_ilu},
        "roofline": {"kernel": "spmv_csr_f64", "bound": "hbm", "achieved": round(kern["spmv_csr_f64_GBps"], 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(kern["spmv_csr_f64_GBps"] / HBM_PEAK_GBS, 4),
                     "traffic": None},
        "cpu_baseline": cpu,
    }
    S.lib().senk_prec_destroy(M)
    S.lib().senk_mat_destroy(A)
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
