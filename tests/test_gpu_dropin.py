"""The C++ drop-in header (include/senk_gpu.hpp) used exactly as a reference user would
call senk::solver::AbmcIluGmresm: a program built with g++ against libsenk_amd.so solves
a system with host arrays, prints the reference's lines, and returns x.  Checked against
the oracle's AbmcIluGmresm: iteration count within 5% (at least +-2) and the FP64 true
relative residual at the tolerance."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import senk_amd as S
from oracle import oracle as O
from tests.helpers import abmc_system, ilu_factors

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parents[1]


def _write(path, N, cnum, bsize, outer, m, sy, L, U, b, eps):
    with open(path, "wb") as f:
        for arr, dt in ((np.array([N, cnum, bsize, outer, m]), np.int32), (sy["val"], np.float64),
                        (sy["cind"], np.int32), (sy["rptr"], np.int32), (L[0], np.float64), (L[1], np.int32),
                        (L[2], np.int32), (U[0], np.float64), (U[1], np.int32), (U[2], np.int32),
                        (sy["cptr"], np.int32), (b, np.float64), (np.array([eps]), np.float64)):
            a = np.ascontiguousarray(arr, dtype=dt)
            f.write(np.int64(a.size).tobytes())
            f.write(a.tobytes())


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_dropin_abmc_ilu_gmresm(tmp_path):
    exe = tmp_path / "driver"
    r = subprocess.run(["g++", "-std=c++17", "-O1", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "dropin" / "driver.cpp"),
                        f"-L{ROOT / 'senk_amd'}", "-lsenk_amd", f"-Wl,-rpath,{ROOT / 'senk_amd'}", "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    rptr, cind, val = S.gen_poisson7(16)
    sy = abmc_system(rptr, cind, val, 4)
    N, eps, m, outer = sy["N"], 1e-8, 30, 200
    xs = S.gen_uniform(1, N)
    b = np.zeros(N)
    O.orc().or_spmv_csr_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), O.P(xs), O.P(b), N)
    L, U = ilu_factors(sy)
    conv = C.c_int()
    xo = np.zeros(N)
    it_o = O.orc().or_gmres_d(O.P(sy["val"]), O.P(sy["cind"]), O.P(sy["rptr"]), C.byref(O.prec_spec(1, L, U)),
                              O.P(b), O.P(xo), float(np.linalg.norm(b)), outer, m, N, eps, None, 0, C.byref(conv))
    inp, out = tmp_path / "sys.bin", tmp_path / "x.bin"
    _write(inp, N, sy["nc"], 4, outer, m, sy, L, U, b, eps)
    run = subprocess.run([str(exe), str(inp), str(out)], capture_output=True, text=True, timeout=120)
    assert run.returncode == 0, run.stdout + run.stderr
    x = np.fromfile(out, dtype=np.float64)
    # the reference prints "I iter <count>" on convergence (ITER_SYMBOL, senk.hpp:11)
    it = re.findall(r"^I iter (\d+)$", run.stdout, re.M)
    assert it, run.stdout
    assert abs(int(it[-1]) - it_o) <= max(2, int(np.ceil(0.05 * it_o))), (it, it_o)
    rows = np.repeat(np.arange(N), np.diff(sy["rptr"]))
    Ax = np.bincount(rows, weights=sy["val"] * x[sy["cind"]], minlength=N)
    assert np.linalg.norm(b - Ax) / np.linalg.norm(b) <= 10 * eps
