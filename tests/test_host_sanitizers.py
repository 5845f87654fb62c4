"""The host-only MatrixMarket reader (senk_amd/csrc/host_io.cpp: mapped file, parallel
token count and parse, counting sorts, Expand, RemoveZeros) built with g++ under the
address and undefined-behaviour sanitizers and run over generated, edge-case and
malformed files, one of them above the parallel-parse threshold: no sanitizer report,
and the statuses the library returns."""
import shutil
import subprocess
from pathlib import Path

import pytest

from tests.helpers import mm_text
from tests.test_matrix_market import BAD

ROOT = Path(__file__).resolve().parents[1]
HERE = Path(__file__).resolve().parent / "sanitize"


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_reader_under_sanitizers(tmp_path):
    # host_io.cpp includes "internal.hpp" from its own directory: build a copy beside the stand-in
    shutil.copy(ROOT / "senk_amd" / "csrc" / "host_io.cpp", tmp_path / "host_io.cpp")
    shutil.copy(HERE / "internal.hpp", tmp_path / "internal.hpp")
    exe = tmp_path / "read"
    r = subprocess.run(["g++", "-O1", "-g", "-std=c++17", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
                        f"-I{ROOT / 'include'}", str(tmp_path / "host_io.cpp"), str(HERE / "read_main.cpp"),
                        "-pthread", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    files, want = [], []
    for name, text in BAD.items():
        files.append(tmp_path / f"bad_{name}.mtx")
        files[-1].write_text(text)
        want.append(13)
    for kind, n, m in (("general", 40, None), ("general", 20, 37), ("symmetric", 50, None), ("integer", 33, 20)):
        files.append(tmp_path / f"{kind}_{n}_{m}.mtx")
        files[-1].write_text(mm_text(kind, 3, n, m))
        want.append(0)
    big = "%%MatrixMarket matrix coordinate real symmetric\n60000 60000 180000\n" + "".join(
        f"{i + 1}\n{i + 1} {2.0 + i % 7}\n" if k == 0 else f"{i + 1} {i + 1 - k}  \n{-0.25 * k}\n"
        for i in range(60000) for k in (0, 1, 2) if i - k >= 0)
    big = big.replace("60000 60000 180000", f"60000 60000 {60000 * 3 - 3}", 1)
    assert len(big) > (1 << 20)
    for name, text, code in (("empty", "%%MatrixMarket matrix coordinate real general\n3 3 0\n", 0), ("zero_bytes", "", 13),
                             ("no_newline", "%%MatrixMarket matrix coordinate real general\n2 2 1\n1 1 5", 0),
                             ("big", big, 0)):
        files.append(tmp_path / f"{name}.mtx")
        files[-1].write_text(text)
        want.append(code)
    run = subprocess.run([str(exe), *map(str, files)], capture_output=True, text=True, timeout=300,
                         env={"ASAN_OPTIONS": "detect_leaks=1:abort_on_error=0", "UBSAN_OPTIONS": "print_stacktrace=1",
                              "PATH": "/usr/bin:/bin"})
    assert "Sanitizer" not in run.stderr and "runtime error" not in run.stderr, run.stderr[-3000:]
    assert run.returncode == 0, run.stderr[-3000:]
    got = [int(line.split()[0]) for line in run.stdout.splitlines()]
    assert got == want, list(zip(map(str, files), got, want))
    last = run.stdout.splitlines()[-1].split()
    assert last == ["0", "0", str(60000 * 3 - 3), "0", "0"]
