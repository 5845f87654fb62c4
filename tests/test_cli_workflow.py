"""The command-line workflow senk_amd/senk_main (the reference's main.cpp:9-187 on the
device) against the reference's own stdout for the same MatrixMarket files
(tests/golden/workflow_*.json, written by tests/golden/make_golden.py from the compiled
reference run step by step).

* CPU: every line the host steps print (reader, structure check, expansion, padding,
  ABMC color sizes, scaling, 8x1 block fill) is identical; the first device step then
  fails loudly (no device, no CPU fallback).
* GPU: the whole output matches, the solve's iteration count within 5% (at least +-2)
  of the reference's, the implicit residual at the tolerance and the FP64 true
  residual ("# test") near the reference's.
"""
import hashlib
import json
import re
import subprocess
from pathlib import Path

import pytest

import senk_amd as S
from senk_amd.build import CLI, LIB, build_cli
from tests.helpers import WORKFLOW_CASES, workflow_text

GOLDEN = Path(__file__).resolve().parent / "golden"


def _case(tmp_path, name):
    g = json.loads((GOLDEN / f"workflow_{name}.json").read_text())
    text = workflow_text(name)
    assert hashlib.sha256(text.encode()).hexdigest() == g["file_sha256"]
    path = tmp_path / f"{name}.mtx"
    path.write_text(text)
    return path, g["stdout"]


def _cli():
    # the workflow program is built with the library; build it alone if only the library is
    if not CLI.exists() and LIB.exists():
        build_cli()
    if not CLI.exists():
        pytest.fail(f"{CLI} is not built (python -m senk_amd.build)")
    return str(CLI)


def _run(path):
    return subprocess.run([_cli(), str(path)], capture_output=True, text=True, timeout=300)


def _has_device():
    try:
        return S.device_count() > 0
    except Exception:
        return False


@pytest.mark.parametrize("name", list(WORKFLOW_CASES))
def test_host_steps_print_the_reference_lines(tmp_path, name):
    if _has_device():
        pytest.skip("a device is present: the full run is checked by the GPU test")
    path, want = _case(tmp_path, name)
    run = _run(path)
    got = run.stdout.splitlines()
    host = want[:want.index("# Allocated!") + 1]
    assert got[:len(host)] == host, run.stdout
    assert run.returncode == 1 and got[len(host)].startswith("senk_ctx_create: no HIP device"), run.stdout


def test_usage_and_bad_file(tmp_path):
    cli = _cli()
    assert subprocess.run([cli], capture_output=True, text=True).returncode == 2
    run = subprocess.run([cli, str(tmp_path / "absent.mtx")], capture_output=True, text=True)
    assert run.returncode == 1 and "file could not be read" in run.stdout
    bad = tmp_path / "bad.mtx"
    bad.write_text("%%MatrixMarket matrix coordinate pattern general\n2 2 1\n1 1\n")
    run = subprocess.run([cli, str(bad)], capture_output=True, text=True)
    assert run.returncode == 1 and "not a supported MatrixMarket matrix" in run.stdout
    nodiag = tmp_path / "nodiag.mtx"
    nodiag.write_text("%%MatrixMarket matrix coordinate real general\n2 2 2\n1 1 1.0\n2 1 1.0\n")
    run = subprocess.run([cli, str(nodiag)], capture_output=True, text=True)
    assert run.returncode == 1 and run.stdout.splitlines()[-1] == "False: CheckStructure"


@pytest.mark.gpu
@pytest.mark.parametrize("name", list(WORKFLOW_CASES))
def test_full_workflow_on_the_device(tmp_path, name):
    path, want = _case(tmp_path, name)
    print(f"workflow {name}", flush=True)
    run = _run(path)
    assert run.returncode == 0, run.stdout + run.stderr
    got = run.stdout.splitlines()
    head = want[:want.index("# Splited!") + 1]
    assert got[:len(head)] == head, run.stdout
    it_ref = int(re.search(r"^I iter (\d+)$", "\n".join(want), re.M).group(1))
    m = re.search(r"^I iter (\d+)$", run.stdout, re.M)
    assert m, run.stdout
    assert abs(int(m.group(1)) - it_ref) <= max(2, int(0.05 * it_ref + 0.999)), (m.group(1), it_ref)
    res = float(re.search(r"^# res (\S+)$", run.stdout, re.M).group(1))
    assert res <= 1e-8
    tail = got[len(head):]
    assert len(tail) == 4 and re.fullmatch(r"\d\.\d{6}e[+-]\d\d", tail[2]), tail  # elapsed ms, %e
    t_ref = float(want[-1].split()[-1])
    t = float(tail[3].split()[-1])
    assert tail[3].startswith("# test ") and t <= max(10 * t_ref, 1e-8), (t, t_ref)
