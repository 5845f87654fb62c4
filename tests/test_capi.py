"""The C ABI library loads and exports every entry point include/senk_amd.h declares;
the C++ drop-in header compiles against it.  CPU only: no device work is started."""
import ctypes as C
import re
import shutil
import subprocess
from pathlib import Path

import pytest

import senk_amd as S

ROOT = Path(__file__).resolve().parents[1]


def test_library_exports_every_declared_symbol():
    L = C.CDLL(str(S.LIB_PATH))
    protos = S.parse_header()
    text = re.sub(r"/\*.*?\*/", "", S.HEADER.read_text(), flags=re.S)
    declared = set(re.findall(r"\b(senk_[a-z0-9_]+)\s*\(", text))
    assert declared == set(protos), declared ^ set(protos)
    missing = [n for n in sorted(declared) if not hasattr(L, n)]
    assert not missing, missing
    assert len(declared) > 100


def test_status_strings_and_version():
    L = S.lib()
    assert L.senk_status_string(0) == b"ok"
    assert b"pivot" in L.senk_status_string(3)
    assert L.senk_version() >= 100


def test_context_without_device_reports_no_device():
    if S.device_count() > 0:
        pytest.skip("a device is visible")
    h = C.c_void_p()
    assert S.lib().senk_ctx_create(0, C.byref(h)) == 7  # SENK_ERR_NO_DEVICE


def test_argument_errors_are_status_codes():
    L = S.lib()
    assert L.senk_spmv_csr_f64(None, None, None, None, None, None, 10) == 1
    assert L.senk_abmc_permutation(None, None, None, None, None, None, 8, 4, 0, b"connect") == 1


def test_ggen_grot_trsv_host_functions():
    c, s = C.c_double(), C.c_double()
    r = S.lib().senk_ggen(3.0, 4.0, C.byref(c), C.byref(s))
    assert r == 5.0 and c.value == 0.6 and s.value == -0.8
    a, b = C.c_double(3.0), C.c_double(4.0)
    S.lib().senk_grot(c.value, s.value, C.byref(a), C.byref(b))
    assert abs(a.value - 5.0) < 1e-15 and abs(b.value) < 1e-15


REF_INCLUDE = Path("/root/reference/include")
needs_gxx = pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")


def _syntax(src, *extra, std="c++17"):
    return subprocess.run(["g++", f"-std={std}", "-fsyntax-only", "-Wall", f"-I{ROOT / 'include'}", *extra, str(src)],
                          capture_output=True, text=True)


@needs_gxx
def test_dropin_header_compiles_every_template():
    """Every drop-in template for double and float, as tests/dropin/instantiate.cpp calls
    them with the reference's argument types."""
    r = _syntax(ROOT / "tests" / "dropin" / "instantiate.cpp")
    assert r.returncode == 0, r.stderr


@needs_gxx
@pytest.mark.skipif(not (REF_INCLUDE / "senk.hpp").exists(), reason="reference headers not present")
def test_dropin_header_beside_reference_headers(tmp_path):
    """Included first, the drop-in takes the solver / kernel / BLAS headers' place and the
    reference's senk.hpp (its host setup code) compiles after it, as the integration notes
    show; included after the reference's headers it is refused with its own #error."""
    r = _syntax(ROOT / "tests" / "dropin" / "instantiate.cpp", "-DSENK_WITH_REFERENCE", f"-I{REF_INCLUDE}", "-fopenmp",
                std="c++20")
    assert r.returncode == 0, r.stderr
    snippet = re.search(r"```cpp\n(#include \"senk_gpu.hpp\".*?)```", (ROOT / "INTEGRATION.md").read_text(), re.S)
    assert snippet, "INTEGRATION.md lost its drop-in snippet"
    src = tmp_path / "snippet.cpp"
    src.write_text(snippet.group(1) + "int main() { return 0; }\n")
    r = _syntax(src, f"-I{REF_INCLUDE}", "-fopenmp", std="c++20")
    assert r.returncode == 0, r.stderr
    bad = tmp_path / "bad.cpp"
    bad.write_text('#include "senk.hpp"\n#include "senk_gpu.hpp"\nint main() { return 0; }\n')
    r = _syntax(bad, f"-I{REF_INCLUDE}", "-fopenmp", std="c++20")
    assert r.returncode != 0 and "include it before senk.hpp" in r.stderr
