"""Python access to libsenk_amd.so, the MI355X (gfx950) implementation of SenK's
preconditioned Krylov path.

The shared library is the product; this module only binds its C ABI
(include/senk_amd.h) with ctypes and adds small conveniences for tests and the
benchmark driver (device arrays, numpy conversion, host-array setup helpers).
Prototypes are read from the header itself, so the binding cannot drift from it.

There is no CPU fallback: importing works without a GPU (so the library's exports can
be inspected), but every call that needs the device fails with SenkError when no HIP
device is present.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

import numpy as np

_PKG = Path(__file__).resolve().parent
LIB_PATH = _PKG / "libsenk_amd.so"
HEADER = _PKG.parent / "include" / "senk_amd.h"

STATUS = {
    0: "SENK_OK", 1: "SENK_ERR_ARG", 2: "SENK_ERR_HIP", 3: "SENK_ERR_ZERO_PIVOT",
    4: "SENK_ERR_DIVISIBILITY", 5: "SENK_ERR_BREAKDOWN", 6: "SENK_ERR_NOT_CONVERGED",
    7: "SENK_ERR_NO_DEVICE", 8: "SENK_ERR_ALLOC", 9: "SENK_ERR_UNSUPPORTED", 10: "SENK_ERR_SCHEDULE",
    11: "SENK_ERR_TIMEOUT", 12: "SENK_ERR_IO", 13: "SENK_ERR_FORMAT",
}
SENK_GMRES, SENK_FGMRES, SENK_GCR, SENK_FGCR, SENK_BICGSTAB, SENK_CG = range(6)
SENK_MGS, SENK_CGS2 = 0, 1
SENK_ILU0, SENK_ILUP, SENK_ILUB, SENK_IC0, SENK_ICP = 0, 1, 2, 3, 4
SENK_FACTOR_F64, SENK_FACTOR_F32, SENK_FACTOR_F16 = 0, 1, 2
SENK_SHAPE_SYM, SENK_SHAPE_UNSYM = 0, 1


class SenkError(RuntimeError):
    def __init__(self, fn: str, status: int, detail: str = ""):
        self.status = status
        super().__init__(f"{fn}: {STATUS.get(status, status)}" + (f" ({detail})" if detail else ""))


class senk_result(C.Structure):
    _fields_ = [("iters", C.c_int), ("converged", C.c_int), ("restarts", C.c_int), ("inner_iters", C.c_int),
                ("fell_back", C.c_int), ("status", C.c_int), ("implicit_relres", C.c_double),
                ("true_relres", C.c_double), ("t_solve_ms", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class senk_sellcs(C.Structure):
    _fields_ = [("N", C.c_int), ("C", C.c_int), ("sigma", C.c_int), ("nslices", C.c_int), ("nnz_padded", C.c_int),
                ("slice_ptr", C.c_void_p), ("cind", C.c_void_p), ("perm", C.c_void_p), ("val64", C.c_void_p),
                ("val32", C.c_void_p)]


class senk_ilu_opts(C.Structure):
    _fields_ = [("kind", C.c_int), ("p", C.c_int), ("bnl", C.c_int), ("bnw", C.c_int), ("precision", C.c_int),
                ("csr_only", C.c_int)]


class senk_solve_opts(C.Structure):
    _fields_ = [("solver", C.c_int), ("m", C.c_int), ("max_iter", C.c_int), ("eps", C.c_double),
                ("ortho", C.c_int), ("mixed", C.c_int), ("factor_prec", C.c_int), ("inner_eps", C.c_double),
                ("inner_max", C.c_int), ("check_every", C.c_int), ("use_learned", C.c_int), ("verbose", C.c_int),
                ("outer_m", C.c_int)]


# ----------------------------------------------------------------- header-driven prototypes
_CTYPES = {"int": C.c_int, "double": C.c_double, "float": C.c_float, "size_t": C.c_size_t, "void": None,
           "unsigned": C.c_uint}


def _ctype(decl: str):
    decl = decl.replace("const ", "").strip()
    if "*" in decl:
        if re.match(r"^char\s*\*$", decl.split("*")[0] + "*") and decl.count("*") == 1 and decl.startswith("char"):
            return C.c_char_p
        return C.c_void_p
    base = decl.split()[0]
    if base in ("senk_ctx", "senk_trsv_plan", "senk_mat", "senk_prec", "senk_gnn"):
        return C.c_void_p
    return _CTYPES[base]


def parse_header(path: Path = HEADER):
    """Return {name: (restype, [argtypes])} for every function declared in the header."""
    text = path.read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    text = re.sub(r"//[^\n]*", "", text)
    protos = {}
    # statements end with ';' (or a block); each declaration is one statement
    for stmt in re.split(r"[;{}]", text):
        stmt = " ".join(stmt.split())
        m = re.match(r"^((?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\*?)\s*(senk_[a-z0-9_]+)\s*\((.*)\)$", stmt)
        if not m or stmt.startswith("typedef"):
            continue
        ret, name, params = m.group(1).strip(), m.group(2), m.group(3).strip()
        args = []
        if params and params != "void":
            for p in params.split(","):
                p = p.strip()
                if p.endswith("*"):
                    t = p
                elif "*" in p:
                    t = p[: p.rfind("*") + 1]
                else:
                    t = re.sub(r"\s+[A-Za-z_][A-Za-z0-9_]*$", "", p)
                args.append(_ctype(t))
        rt = C.c_char_p if ret.replace(" ", "") in ("constchar*", "char*") else _ctype(ret)
        protos[name] = (rt, args)
    return protos


_lib = None


def lib():
    """The loaded library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m senk_amd.build`")
        L = C.CDLL(str(LIB_PATH))
        for name, (rt, args) in parse_header().items():
            fn = getattr(L, name)
            fn.restype = rt
            fn.argtypes = args
        _lib = L
    return _lib


def check(fn: str, status: int):
    if status != 0:
        detail = lib().senk_last_error().decode() if status in (2, 12, 13) else ""
        raise SenkError(fn, status, detail)


def call(name: str, *args):
    """Call a status-returning entry point and raise on error."""
    st = getattr(lib(), name)(*args)
    check(name, st)
    return st


def device_count() -> int:
    n = C.c_int(0)
    lib().senk_device_count(C.byref(n))
    return n.value


# ----------------------------------------------------------------- device memory
_DT = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32"}


class DeviceArray:
    """A 1-D device buffer owned by a Context."""

    def __init__(self, ctx: "Context", n: int, dtype):
        self.ctx = ctx
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        p = C.c_void_p()
        call("senk_device_malloc", ctx.h, C.c_size_t(max(1, self.n * self.dtype.itemsize)), C.byref(p))
        self.ptr = p.value

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, a: np.ndarray):
        a = np.ascontiguousarray(a, dtype=self.dtype)
        assert a.size == self.n
        call("senk_memcpy_htod", self.ctx.h, C.c_void_p(self.ptr), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes))
        return self

    def download(self) -> np.ndarray:
        out = np.empty(self.n, dtype=self.dtype)
        call("senk_memcpy_dtoh", self.ctx.h, out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), C.c_size_t(out.nbytes))
        return out

    def zero(self):
        call("senk_memset_zero", self.ctx.h, C.c_void_p(self.ptr), C.c_size_t(self.nbytes))
        return self

    def free(self):
        if self.ptr:
            lib().senk_device_free(self.ctx.h, C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            if self.ptr and self.ctx.h:
                self.free()
        except Exception:
            pass

    @property
    def p(self):
        return C.c_void_p(self.ptr)


class Context:
    def __init__(self, device: int = 0):
        h = C.c_void_p()
        call("senk_ctx_create", C.c_int(device), C.byref(h))
        self.h = h

    def array(self, a) -> DeviceArray:
        a = np.ascontiguousarray(a)
        return DeviceArray(self, a.size, a.dtype).upload(a)

    def empty(self, n, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, n, dtype)

    def zeros(self, n, dtype=np.float64) -> DeviceArray:
        return DeviceArray(self, n, dtype).zero()

    def sync(self):
        call("senk_ctx_synchronize", self.h)

    def close(self):
        if self.h:
            lib().senk_ctx_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


# ----------------------------------------------------------------- host-array helpers
def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _take(ptr: C.c_void_p, n: int, dtype) -> np.ndarray:
    """Copy a malloc'd library output into numpy and free it."""
    dtype = np.dtype(dtype)
    if n == 0:
        lib().senk_host_free(ptr)
        return np.empty(0, dtype=dtype)
    buf = (C.c_char * (n * dtype.itemsize)).from_address(ptr.value)
    out = np.frombuffer(bytes(buf), dtype=dtype).copy()
    lib().senk_host_free(ptr)
    return out


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _csr_out(fn: str, *args):
    N = C.c_int()
    rp, ci, v = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call(fn, *args, C.byref(N), C.byref(rp), C.byref(ci), C.byref(v))
    n = N.value
    rptr = _take(rp, n + 1, np.int32)
    nnz = int(rptr[-1])
    return rptr, _take(ci, nnz, np.int32), _take(v, nnz, np.float64)


def gen_poisson7(nx, ny=None, nz=None):
    ny = nx if ny is None else ny
    nz = nx if nz is None else nz
    return _csr_out("senk_gen_poisson7", nx, ny, nz)


def gen_convdiff27(n, peclet=10.0):
    return _csr_out("senk_gen_convdiff27", n, C.c_double(peclet))


def gen_vardiff7(n, seed=1, corr_len=8.0, sigma=1.5):
    return _csr_out("senk_gen_vardiff7", n, seed, C.c_double(corr_len), C.c_double(sigma))


def gen_unstructured(n, seed=1):
    return _csr_out("senk_gen_unstructured", n, seed)


def gen_elasticity3(n, seed=1, nu=0.3):
    return _csr_out("senk_gen_elasticity3", n, seed, C.c_double(nu))


def gen_uniform(seed, n):
    out = np.empty(n, dtype=np.float64)
    call("senk_gen_uniform", seed, n, _ptr(out))
    return out


def check_structure(cind, rptr, N):
    return bool(lib().senk_check_structure(_ptr(_i32(cind)), _ptr(_i32(rptr)), N))


def abmc_permutation(cind, rptr, N, bsize, isSym=False, method="connect"):
    nc = C.c_int()
    sc, lp, rp = C.c_void_p(), C.c_void_p(), C.c_void_p()
    cind, rptr = _i32(cind), _i32(rptr)
    call("senk_abmc_permutation", _ptr(cind), _ptr(rptr), C.byref(nc), C.byref(sc), C.byref(lp), C.byref(rp), N,
         bsize, int(isSym), method.encode())
    return nc.value, _take(sc, nc.value + 1, np.int32), _take(lp, N, np.int32), _take(rp, N, np.int32)


def amc_permutation(cind, rptr, N, isSym=False):
    nc = C.c_int()
    sc, lp, rp = C.c_void_p(), C.c_void_p(), C.c_void_p()
    cind, rptr = _i32(cind), _i32(rptr)
    call("senk_amc_permutation", _ptr(cind), _ptr(rptr), C.byref(nc), C.byref(sc), C.byref(lp), C.byref(rp), N,
         int(isSym))
    return nc.value, _take(sc, nc.value + 1, np.int32), _take(lp, N, np.int32), _take(rp, N, np.int32)


def reordering(val, cind, rptr, LP, RP, N):
    """B = P A P^T; returns new (val, cind, rptr) (inputs untouched)."""
    val, cind, rptr = _f64(val).copy(), _i32(cind).copy(), _i32(rptr).copy()
    call("senk_reordering", _ptr(val), _ptr(cind), _ptr(rptr), _ptr(_i32(LP)), _ptr(_i32(RP)), N)
    return val, cind, rptr


def reordering_rp(val, cind, rptr, RP, N):
    val, cind, rptr = _f64(val).copy(), _i32(cind).copy(), _i32(rptr)
    call("senk_reordering_rp", _ptr(val), _ptr(cind), _ptr(rptr), _ptr(_i32(RP)), N)
    return val, cind


_SPLIT_KEYS = {"LD-U": 0, "L-DU": 1, "L-D-U": 2, "LU-D": 3}


def split(val, cind, rptr, N, key="L-DU", invDiag=True):
    """Returns dict with L (and U, diag as the key dictates) as numpy arrays."""
    ptrs = [C.c_void_p() for _ in range(7)]
    val, cind, rptr = _f64(val), _i32(cind), _i32(rptr)
    k = _SPLIT_KEYS.get(key, -1)
    call("senk_split", _ptr(val), _ptr(cind), _ptr(rptr), *(C.byref(p) for p in ptrs[:6]), C.byref(ptrs[6]), N,
         key.encode(), int(invDiag))
    out = {}
    lrptr = _take(ptrs[2], N + 1, np.int32)
    out["L"] = (_take(ptrs[0], int(lrptr[-1]), np.float64), _take(ptrs[1], int(lrptr[-1]), np.int32), lrptr)
    if k != 3:
        urptr = _take(ptrs[5], N + 1, np.int32)
        out["U"] = (_take(ptrs[3], int(urptr[-1]), np.float64), _take(ptrs[4], int(urptr[-1]), np.int32), urptr)
    if k >= 2:
        out["diag"] = _take(ptrs[6], N, np.float64)
    return out


def _pattern_inplace(fn, val, cind, rptr, N, *extra):
    # hand over malloc'd copies (the library frees and replaces them, as the reference does)
    pv, pc, pr = _malloc_copy(_f64(val)), _malloc_copy(_i32(cind)), _malloc_copy(_i32(rptr))
    call(fn, C.byref(pv), C.byref(pc), C.byref(pr), N, *extra)
    rp = _take(pr, N + 1, np.int32)
    nnz = int(rp[-1])
    return _take(pv, nnz, np.float64), _take(pc, nnz, np.int32), rp


def alloc_level_zero(val, cind, rptr, N, p):
    return _pattern_inplace("senk_alloc_level_zero", val, cind, rptr, N, p)


def alloc_block_zero(val, cind, rptr, N, bnl, bnw):
    return _pattern_inplace("senk_alloc_block_zero", val, cind, rptr, N, bnl, bnw)


def remove_off_diagonal(val, cind, rptr, N, bnum):
    return _pattern_inplace("senk_remove_off_diagonal", val, cind, rptr, N, bnum)


def csr2bcsr(val, cind, rptr, N, bnl, bnw):
    bv, bc, br = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call("senk_csr2bcsr", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), C.byref(bv), C.byref(bc), C.byref(br),
         N, bnl, bnw)
    brptr = _take(br, N // bnl + 1, np.int32)
    nb = int(brptr[-1])
    return _take(bv, nb * bnl * bnw, np.float64), _take(bc, nb, np.int32), brptr


def bcsr2csr(bval, bcind, brptr, N, bnl, bnw):
    v, c, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int()
    call("senk_bcsr2csr", _ptr(_f64(bval)), _ptr(_i32(bcind)), _ptr(_i32(brptr)), C.byref(v), C.byref(c), C.byref(r),
         N, bnl, bnw, C.byref(nnz))
    return _take(v, nnz.value, np.float64), _take(c, nnz.value, np.int32), _take(r, N + 1, np.int32)


def csr2sell(val, cind, rptr, size, N):
    v, c, l = C.c_void_p(), C.c_void_p(), C.c_void_p()
    nnz = C.c_int()
    call("senk_csr2sell", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), C.byref(v), C.byref(c), C.byref(l),
         size, N, C.byref(nnz))
    ns = (N + size - 1) // size
    return _take(v, nnz.value, np.float64), _take(c, nnz.value, np.int32), _take(l, ns + 1, np.int32)


def csr2sellcs(val, cind, rptr, N, C_=64, sigma=1024):
    s = senk_sellcs()
    call("senk_csr2sellcs", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), N, C_, sigma, C.byref(s))
    return s


def _malloc_copy(a: np.ndarray) -> C.c_void_p:
    m = C.CDLL(None).malloc
    m.restype = C.c_void_p
    m.argtypes = [C.c_size_t]
    p = C.c_void_p(m(max(1, a.nbytes)))
    C.memmove(p.value, a.ctypes.data, a.nbytes)
    return p


def padding(val, cind, rptr, size, N):
    """Padding (identity rows up to a multiple of size); returns (val, cind, rptr, N)."""
    pv, pc, pr = _malloc_copy(_f64(val)), _malloc_copy(_i32(cind)), _malloc_copy(_i32(rptr))
    n = C.c_int(N)
    call("senk_padding", C.byref(pv), C.byref(pc), C.byref(pr), size, C.byref(n))
    r = _take(pr, n.value + 1, np.int32)
    nnz = int(r[-1])
    return _take(pv, nnz, np.float64), _take(pc, nnz, np.int32), r, n.value


def csr2csc(val, cind, rptr, N, M):
    v, r, c = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call("senk_csr2csc", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), C.byref(v), C.byref(r), C.byref(c), N, M)
    nnz = int(np.asarray(rptr)[N])
    return _take(v, nnz, np.float64), _take(r, nnz, np.int32), _take(c, M + 1, np.int32)


def read_matrix_market(path, remove_zeros=True):
    """io::ReadMatrixMarket (senk_io.hpp:41-103): (val, cind, rptr, N, M, shape), shape
    SENK_SHAPE_SYM (the stored lower triangle; complete it with expand) or SENK_SHAPE_UNSYM."""
    v, c, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
    N, M, sh = C.c_int(), C.c_int(), C.c_int()
    call("senk_read_matrix_market", os.fsencode(str(path)), C.byref(v), C.byref(c), C.byref(r), C.byref(N), C.byref(M),
         C.byref(sh), int(bool(remove_zeros)))
    rptr = _take(r, N.value + 1, np.int32)
    nnz = int(rptr[-1])
    return _take(v, nnz, np.float64), _take(c, nnz, np.int32), rptr, N.value, M.value, sh.value


def remove_zeros(val, cind, rptr, N):
    """RemoveZeros (senk_matrix.hpp:29-43); returns (val, cind, rptr)."""
    pv, pc = _malloc_copy(_f64(val)), _malloc_copy(_i32(cind))
    rp = _i32(rptr).copy()
    call("senk_remove_zeros", C.byref(pv), C.byref(pc), _ptr(rp), N)
    nnz = int(rp[N])
    return _take(pv, nnz, np.float64), _take(pc, nnz, np.int32), rp


def expand(val, cind, rptr, N, key="L"):
    """Expand (senk_matrix.hpp:335-371): the full matrix from a stored triangle."""
    v, c, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call("senk_expand", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), C.byref(v), C.byref(c), C.byref(r), N,
         key.encode())
    rp = _take(r, N + 1, np.int32)
    nnz = int(rp[-1])
    return _take(v, nnz, np.float64), _take(c, nnz, np.int32), rp


def duplicate(val, cind, rptr, N):
    """Duplicate (senk_matrix.hpp:373-389)."""
    v, c, r = C.c_void_p(), C.c_void_p(), C.c_void_p()
    call("senk_duplicate", _ptr(_f64(val)), _ptr(_i32(cind)), _ptr(_i32(rptr)), C.byref(v), C.byref(c), C.byref(r), N)
    rp = _take(r, N + 1, np.int32)
    nnz = int(rp[-1])
    return _take(v, nnz, np.float64), _take(c, nnz, np.int32), rp


def solve_opts(**kw) -> senk_solve_opts:
    o = senk_solve_opts()
    lib().senk_solve_opts_default(C.byref(o))
    for k, v in kw.items():
        setattr(o, k, v)
    return o


__all__ = [n for n in dir() if not n.startswith("_")]
