"""Writes the golden fixtures in tests/golden/ from the reference itself.

Run in a container that has the reference tree (it needs oracle/_ref/libsenkref.so,
built by `make -C oracle`).  Inputs come from the library's own generators; every
expected output comes from the compiled reference headers.  Only the resulting data
(.npz / .json) is committed, never reference source.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import ctypes as C
import hashlib
import json
import os
import sys
import tempfile
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

import senk_amd as S  # noqa: E402
from oracle import oracle as O  # noqa: E402

OUT = Path(__file__).resolve().parent
P = O.P


def capture(fn):
    """Run fn with the process stdout redirected to a file; return the text."""
    sys.stdout.flush()
    with tempfile.TemporaryFile(mode="w+b") as tmp:
        saved = os.dup(1)
        os.dup2(tmp.fileno(), 1)
        try:
            fn()
            O.ref().ref_flush()
        finally:
            os.dup2(saved, 1)
            os.close(saved)
        tmp.seek(0)
        return tmp.read().decode()


def last_iter(text):
    for line in text.splitlines()[::-1]:
        if line.startswith("I iter ") or line.startswith("# iter "):
            return int(line.split()[2])
    raise RuntimeError(text[-400:])


def sha(*arrays):
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def small_case(name, rptr, cind, val, bsizes, ps):
    ref = O.ref()
    N = len(rptr) - 1
    rng = np.random.default_rng(7)
    x = rng.uniform(-1, 1, N)
    y = np.zeros(N)
    ref.ref_spmv_csr_d(P(val), P(cind), P(rptr), P(x), P(y), N)
    f = val.copy()
    ref.ref_ilu0_d(P(f), P(cind), P(rptr), N)
    nnz = len(val)
    out = [np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(nnz), np.zeros(nnz, np.int32),
           np.zeros(N + 1, np.int32), np.zeros(N)]
    ref.ref_split_d(P(f), P(cind), P(rptr), *(P(t) for t in out), N, 1, 1)
    lv, lc, lr, uv, uc, ur = out[:6]
    yl, yu = np.zeros(N), np.zeros(N)
    ref.ref_sptrsv_l_d(P(lv), P(lc), P(lr), P(x), P(yl), N)
    ref.ref_sptrsv_u_d(P(uv), P(uc), P(ur), P(yl), P(yu), N)
    fx = dict(rptr=rptr, cind=cind, val=val, x=x, spmv=y, ilu0=f, ilu_apply=yu)
    for bsize in bsizes:
        for method in (0, 1):
            sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
            nc = ref.ref_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, bsize, 0, method)
            fx[f"abmc_{bsize}_{method}"] = np.concatenate([[nc], sc[: nc + 1], lp]).astype(np.int32)
    for p in ps:
        cap = nnz * 40
        ov, oc, orr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
        n = ref.ref_alloc_level_zero_d(P(val), P(cind), P(rptr), P(ov), P(oc), P(orr), N, p, cap)
        fx[f"ilup_{p}"] = np.concatenate([orr, oc[:n]]).astype(np.int32)
    b = S.gen_uniform(1, N)
    xr = np.zeros(N)
    txt = capture(lambda: ref.ref_ilu_gmresm_d(P(val), P(cind), P(rptr), P(lv), P(lc), P(lr), P(uv), P(uc), P(ur),
                                               P(b), P(xr), float(np.linalg.norm(b)), 100, 30, N, 1e-8))
    fx["b"] = b
    fx["gmres_iters"] = np.array(last_iter(txt))
    np.savez_compressed(OUT / f"{name}.npz", **fx)
    print(f"wrote {name}.npz")


def poisson32_case():
    """7-point Poisson 32^3, ABMC bsize 4 "connect", ILU(0), AbmcIluGmresm m=30 to
    1e-8 with b = A x*, x* ~ U[-1,1] (seed 1)."""
    ref = O.ref()
    rptr, cind, val = S.gen_poisson7(32)
    N = len(rptr) - 1
    sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc = ref.ref_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, 4, 0, 1)
    v, c, r = val.copy(), cind.copy(), rptr.copy()
    ref.ref_reordering_d(P(v), P(c), P(r), P(lp), P(rp), N)
    f = v.copy()
    ref.ref_ilu0_d(P(f), P(c), P(r), N)
    nnz = len(v)
    out = [np.zeros(nnz), np.zeros(nnz, np.int32), np.zeros(N + 1, np.int32), np.zeros(nnz), np.zeros(nnz, np.int32),
           np.zeros(N + 1, np.int32), np.zeros(N)]
    ref.ref_split_d(P(f), P(c), P(r), *(P(t) for t in out), N, 1, 1)
    xs = S.gen_uniform(1, N)
    b = np.zeros(N)
    ref.ref_spmv_csr_d(P(val), P(cind), P(rptr), P(xs), P(b), N)
    bp = b[lp].copy()
    x = np.zeros(N)
    txt = capture(lambda: ref.ref_abmc_ilu_gmresm_d(P(v), P(c), P(r), *(P(t) for t in out[:6]), P(sc), nc, 4, P(bp),
                                                    P(x), float(np.linalg.norm(bp)), 200, 30, N, 1e-8))
    it = last_iter(txt)
    res_line = [ln for ln in txt.splitlines() if ln.startswith("# res")][-1]
    t = np.zeros(N)
    ref.ref_spmv_csr_d(P(v), P(c), P(r), P(x), P(t), N)
    true_rel = float(np.linalg.norm(bp - t) / np.linalg.norm(bp))
    np.savez_compressed(OUT / "poisson7_32_abmc4.npz", LP=lp, size_color=sc[: nc + 1], num_color=np.array(nc),
                        gmres_iters=np.array(it))
    meta = {"config": "poisson7 32^3, ABMC bsize 4 connect, ILU(0), AbmcIluGmresm m=30 eps=1e-8, b=A x*",
            "num_color": int(nc), "gmres_iters": it, "res_line": res_line, "true_relres": true_rel,
            "sha256_LP": sha(lp), "sha256_ilu0": sha(f), "sha256_size_color": sha(sc[: nc + 1])}
    (OUT / "poisson7_32_abmc4.json").write_text(json.dumps(meta, indent=1) + "\n")
    print("wrote c1", meta)


def digest_case(name, gen, bsize, isSym=0, ilup=None):
    """Digests of the reference's ordering arrays at a full benchmark size."""
    ref = O.ref()
    rptr, cind, val = gen()
    N = len(rptr) - 1
    sc, lp, rp = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc = ref.ref_abmc_perm(P(cind), P(rptr), P(sc), P(lp), P(rp), N, bsize, isSym, 1)
    meta = {"N": N, "nnz": int(rptr[-1]), "bsize": bsize, "isSym": isSym, "num_color": int(nc),
            "sha256_LP": sha(lp), "sha256_RP": sha(rp), "sha256_size_color": sha(sc[: nc + 1]),
            "sha256_matrix": sha(rptr, cind, val)}
    if ilup:
        v, c, r = val.copy(), cind.copy(), rptr.copy()
        ref.ref_reordering_d(P(v), P(c), P(r), P(lp), P(rp), N)
        cap = len(v) * 3
        ov, oc, orr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
        n = ref.ref_alloc_level_zero_d(P(v), P(c), P(r), P(ov), P(oc), P(orr), N, ilup, cap)
        meta[f"ilu{ilup}_nnz"] = int(n)
        meta[f"sha256_ilu{ilup}_pattern"] = sha(orr, oc[:n])
    (OUT / f"{name}.json").write_text(json.dumps(meta, indent=1) + "\n")
    print("wrote", name, meta)


MM_CASES = {  # name: (kind, seed, n, m)
    "general": ("general", 11, 36, None), "rect_wide": ("general", 12, 24, 31),
    "rect_tall": ("general", 13, 33, 21), "symmetric": ("symmetric", 14, 40, None),
    "integer": ("integer", 15, 30, None),
}


def _ref_read(path, rz, cap=4096):
    ref = O.ref()
    v, c, r, d = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(3, np.int32)
    nnz = capture_value(lambda: ref.ref_read_mm(str(path).encode(), rz, P(v), P(c), P(r), P(d), cap, cap))
    assert nnz >= 0, nnz
    return v[:nnz].copy(), c[:nnz].copy(), r[:d[0] + 1].copy(), d.copy()


def capture_value(fn):
    out = []
    capture(lambda: out.append(fn()))
    return out[0]


def matrix_market_case():
    """MatrixMarket files made by tests/helpers.mm_text, read by the reference's
    io::ReadMatrixMarket with and without zero removal; its RemoveZeros on the kept-zeros
    read, and its Expand ("L" on the symmetric file's triangle, "U" on the transpose)."""
    from tests.helpers import mm_text
    ref = O.ref()
    fx = {}
    with tempfile.TemporaryDirectory() as d:
        for name, (kind, seed, n, m) in MM_CASES.items():
            text = mm_text(kind, seed, n, m)
            path = Path(d) / f"{name}.mtx"
            path.write_text(text)
            fx[f"{name}_text"] = np.frombuffer(text.encode(), dtype=np.uint8)
            for rz in (1, 0):
                v, c, r, dims = _ref_read(path, rz)
                fx.update({f"{name}_rz{rz}_val": v, f"{name}_rz{rz}_cind": c, f"{name}_rz{rz}_rptr": r,
                           f"{name}_rz{rz}_dims": dims})
            v, c, r, dims = _ref_read(path, 0)
            N = int(dims[0])
            rr = r.copy()
            nnz = ref.ref_remove_zeros_d(P(v), P(c), P(rr), N)
            fx.update({f"{name}_rmz_val": v[:nnz].copy(), f"{name}_rmz_cind": c[:nnz].copy(), f"{name}_rmz_rptr": rr})
            if kind == "symmetric":
                v, c, r, dims = _ref_read(path, 1)
                cap = 2 * len(v)
                for key, (tv, tc, tr) in (("L", (v, c, r)), ("U", tuple(S.csr2csc(v, c, r, N, N)[i] for i in (0, 1, 2)))):
                    ev, ec, er = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
                    k = ref.ref_expand_d(P(tv), P(tc), P(tr), N, key.encode(), P(ev), P(ec), P(er), cap)
                    assert k >= 0
                    fx.update({f"{name}_expand{key}_val": ev[:k].copy(), f"{name}_expand{key}_cind": ec[:k].copy(),
                               f"{name}_expand{key}_rptr": er})
    np.savez_compressed(OUT / "matrix_market.npz", **fx)


def workflow_case(name):
    """The reference's end-to-end workflow (main.cpp:9-187) on a MatrixMarket file, run
    step by step through the compiled reference; its stdout lines (the solve's elapsed
    time replaced by "<elapsed>") and the file's digest go to workflow_<name>.json."""
    from tests.helpers import workflow_text
    ref = O.ref()
    text = workflow_text(name)
    out = []
    with tempfile.TemporaryDirectory() as d:
        path = Path(d) / f"{name}.mtx"
        path.write_text(text)
        cap = 1 << 22
        v, c, r, dims = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(cap, np.int32), np.zeros(3, np.int32)
        res = []
        out.append(capture(lambda: res.append(ref.ref_read_mm(str(path).encode(), 1, P(v), P(c), P(r), P(dims), cap,
                                                              cap))))
    assert res[0] >= 0
    N = int(dims[0])
    v, c, r = v[:res[0]].copy(), c[:res[0]].copy(), r[:N + 1].copy()
    out.append("# %d\n" % r[N])
    assert ref.ref_check_structure_d(P(v), P(c), P(r), N) == 1
    if int(dims[2]) == 0:
        cap = 2 * len(v)
        ev, ec, er = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
        k = ref.ref_expand_d(P(v), P(c), P(r), N, b"L", P(ev), P(ec), P(er), cap)
        v, c, r = ev[:k].copy(), ec[:k].copy(), er
    ori_N = N
    Np = (N + 127) // 128 * 128
    pv, pc, pr = np.zeros(len(v) + Np - N), np.zeros(len(v) + Np - N, np.int32), np.zeros(Np + 1, np.int32)
    pv[:len(v)], pc[:len(v)], pr[:N + 1] = v, c, r
    N = ref.ref_padding_d(P(pv), P(pc), P(pr), 128, N)
    v, c, r = pv, pc, pr
    out.append("%d\n" % r[N])
    out.append("%d %d\n" % (N, N))
    sc, LP, RP = np.zeros(N + 1, np.int32), np.zeros(N, np.int32), np.zeros(N, np.int32)
    nc = ref.ref_abmc_perm(P(c), P(r), P(sc), P(LP), P(RP), N, 128, 0, 1)
    ref.ref_reordering_d(P(v), P(c), P(r), P(LP), P(RP), N)
    out.append("# " + "".join("%d " % (sc[i + 1] - sc[i]) for i in range(nc)) + "\n")
    out.append("# Reordered!\n")
    dg = np.zeros(N)
    ref.ref_get_diag_d(P(v), P(c), P(r), P(dg), N)
    dg = 1 / np.sqrt(np.abs(dg))
    ref.ref_scaling_d(P(v), P(c), P(r), P(dg), P(dg), N)
    cap = 8 * len(v)
    fv, fc, fr = np.zeros(cap), np.zeros(cap, np.int32), np.zeros(N + 1, np.int32)
    fn = ref.ref_alloc_block_zero_d(P(v), P(c), P(r), P(fv), P(fc), P(fr), N, 8, 1, cap)
    assert fn > 0
    fv, fc = fv[:fn].copy(), fc[:fn].copy()
    out.append("# Allocated!\n")
    ref.ref_ilu0_d(P(fv), P(fc), P(fr), N)
    out.append("# Factrized!\n")
    lv, lc, lr = np.zeros(fn), np.zeros(fn, np.int32), np.zeros(N + 1, np.int32)
    uv, uc, ur = np.zeros(fn), np.zeros(fn, np.int32), np.zeros(N + 1, np.int32)
    ref.ref_split_d(P(fv), P(fc), P(fr), P(lv), P(lc), P(lr), P(uv), P(uc), P(ur), None, N, 1, 1)
    out.append("# Splited!\n")
    b = S.gen_uniform(1, N)  # std::mt19937(1), uniform_real_distribution<>(-1, 1)
    b[ori_N:] = 0.0
    x = np.zeros(N)
    nrm_b = ref.ref_nrm2_d(P(b), N)
    out.append(capture(lambda: ref.ref_ilu_gmresm_d(P(v), P(c), P(r), P(lv), P(lc), P(lr), P(uv), P(uc), P(ur),
                                                    P(b), P(x), nrm_b, 6000 // 50, 50, N, 1e-8)))
    out.append("<elapsed>\n")
    out.append(capture(lambda: ref.ref_relative_residual_error(P(v), P(c), P(r), P(b), P(x), N)))
    lines = "".join(out).splitlines()
    doc = {"file_sha256": hashlib.sha256(text.encode()).hexdigest(), "stdout": lines}
    (OUT / f"workflow_{name}.json").write_text(json.dumps(doc, indent=1) + "\n")


def main():
    if O.ref() is None:
        sys.exit("oracle/_ref/libsenkref.so is not built (make -C oracle with the reference tree present)")
    small_case("poisson7_6", *S.gen_poisson7(6), bsizes=(4, 8), ps=(1, 2))
    small_case("convdiff27_5", *S.gen_convdiff27(5, 10.0), bsizes=(5, 25), ps=(1,))
    small_case("unstructured_300", *S.gen_unstructured(300, 4), bsizes=(4, 10), ps=(1, 2))
    small_case("elasticity_3", *S.gen_elasticity3(3, 4), bsizes=(3, 9), ps=(1,))
    small_case("vardiff7_6", *S.gen_vardiff7(6, 3, 4.0, 1.5), bsizes=(4, 12), ps=(1,))
    poisson32_case()
    matrix_market_case()
    from tests.helpers import WORKFLOW_CASES
    for name in WORKFLOW_CASES:
        workflow_case(name)
    if "--full" in sys.argv:
        digest_case("convdiff27_128_abmc16", lambda: S.gen_convdiff27(128, 10.0), 16)
        digest_case("vardiff7_128_abmc16", lambda: S.gen_vardiff7(128, 1, 8.0, 1.5), 16)
        digest_case("elasticity3_96_abmc48", lambda: S.gen_elasticity3(96, 1), 48, ilup=1)
        digest_case("unstructured_4194304_abmc16", lambda: S.gen_unstructured(4194304, 1), 16)


if __name__ == "__main__":
    main()
