"""TEST INFRASTRUCTURE ONLY: ctypes loaders for the CPU oracle.

- ``orc()``  -> oracle/liboracle.so, the plain-C restatement of the reference
  (senk_oracle.c); prototypes are read from senk_oracle.h.
- ``ref()``  -> oracle/_ref/libsenkref.so, the reference's own headers compiled by
  oracle/Makefile (None when it was not built, e.g. without the reference tree).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may use this
module; it is the checker and the CPU baseline, never the product path.
"""
from __future__ import annotations

import ctypes as C
import re
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libsenkref.so"

_T = {"int": C.c_int, "double": C.c_double, "float": C.c_float, "void": None}


def _parse(path: Path):
    text = re.sub(r"/\*.*?\*/", "", path.read_text(), flags=re.S)
    out = {}
    for stmt in re.split(r"[;{}]", text):
        stmt = " ".join(stmt.split())
        m = re.match(r"^([A-Za-z_][A-Za-z0-9_]*\s*\*?)\s*(or_[a-z0-9_]+)\s*\((.*)\)$", stmt)
        if not m:
            continue
        ret, name, params = m.groups()
        args = []
        for p in params.split(","):
            p = p.strip().replace("const ", "")
            args.append(C.c_void_p if "*" in p else _T[p.split()[0]])
        out[name] = (C.c_void_p if "*" in ret else _T[ret.strip()], args)
    return out


_orc = None
_ref = None
_ref_tried = False


def orc():
    global _orc
    if _orc is None:
        if not ORACLE_SO.exists():
            raise ImportError(f"{ORACLE_SO} missing: run `make -C oracle`")
        L = C.CDLL(str(ORACLE_SO))
        for name, (rt, args) in _parse(HERE / "senk_oracle.h").items():
            f = getattr(L, name)
            f.restype, f.argtypes = rt, args
        _orc = L
    return _orc


# Prototypes of the compiled reference shim (ref_shim.cpp): return type, then argument
# kinds: p = pointer, i = int, d = double, f = float.
_REF_PROTOS = {
    "ref_spmv_csr_d": (None, "pppppi"), "ref_spmv_csr_f": (None, "pppppi"), "ref_spmv_csr_diag_d": (None, "ppppppi"),
    "ref_spmv_sell_d": (None, "pppippi"), "ref_spmv_bcsr_d": ("i", "pppppiii"),
    "ref_spmv_sell_f": (None, "pppippi"), "ref_spmv_bcsr_f": ("i", "pppppiii"),
    "ref_sptrsv_l_d": (None, "pppppi"), "ref_sptrsv_u_d": (None, "pppppi"),
    "ref_sptrsv_l_f": (None, "pppppi"), "ref_sptrsv_u_f": (None, "pppppi"),
    "ref_sptrsv_l_amc_d": (None, "pppppipi"), "ref_sptrsv_u_amc_d": (None, "pppppipi"),
    "ref_sptrsv_l_abmc_d": (None, "pppppipii"), "ref_sptrsv_u_abmc_d": (None, "pppppipii"),
    "ref_sptrsv_l_bj_d": (None, "pppppii"), "ref_sptrsv_u_bj_d": (None, "pppppii"),
    "ref_sptrsv_bcsr_l_d": ("i", "pppppiii"), "ref_sptrsv_bcsr_u_d": ("i", "pppppiii"),
    "ref_sptrsv_bcsr_l_abmc_d": ("i", "pppppipiiii"), "ref_sptrsv_bcsr_u_abmc_d": ("i", "pppppipiiii"),
    "ref_copy_d": (None, "ppi"), "ref_scal_d": (None, "dpi"), "ref_axpy_d": (None, "dppi"),
    "ref_axpby_d": (None, "dpdpi"), "ref_axpyz_d": (None, "dpppi"), "ref_dot_d": ("d", "ppi"), "ref_nrm2_d": ("d", "pi"),
    "ref_dot_f": ("f", "ppi"), "ref_nrm2_f": ("f", "pi"), "ref_hadprod_d": (None, "ppi"), "ref_haddiv_d": (None, "ppi"),
    "ref_ggen_d": ("d", "ddpp"), "ref_grot_d": (None, "ddpp"), "ref_trsv_d": (None, "pppii"),
    "ref_convert_d2f": (None, "ppi"),
    "ref_check_structure_d": ("i", "pppi"), "ref_csr2csc_d": (None, "ppppppii"),
    "ref_csr2sell_d": ("i", "ppppppiii"), "ref_csr2bcsr_d": ("i", "ppppppiiii"), "ref_bcsr2csr_d": ("i", "ppppppiiii"),
    "ref_padding_d": ("i", "pppii"), "ref_split_d": (None, "pppppppppp" + "iii"), "ref_get_diag_d": (None, "ppppi"),
    "ref_scaling_d": (None, "pppppi"), "ref_ilu0_d": (None, "pppi"), "ref_ilu0_f": (None, "pppi"),
    "ref_ilup_d": ("i", "ppppppiii"), "ref_alloc_level_zero_d": ("i", "ppppppiii"),
    "ref_alloc_block_zero_d": ("i", "ppppppiiii"), "ref_remove_off_diagonal_d": ("i", "pppii"),
    "ref_reordering_rp_d": (None, "ppppi"), "ref_reordering_d": (None, "pppppi"),
    "ref_amc_perm": ("i", "pppppii"), "ref_abmc_perm": ("i", "pppppiiii"),
    "ref_gmresm_d": (None, "pppppdiiid"), "ref_ilu_gmresm_d": (None, "pppppppppppdiiid"),
    "ref_ilu_gmresm_f": (None, "pppppppppppfiiif"),
    "ref_amc_ilu_gmresm_d": (None, "pppppppppp" + "i" + "ppdiiid"),
    "ref_abmc_ilu_gmresm_d": (None, "ppppppppppii" + "ppdiiid"),
    "ref_bj_ilu_gmresm_d": (None, "pppppppppi" + "ppdiiid"),
    "ref_ilub_gmresm_d": ("i", "ppppppppp" + "ppdiiid" + "ii"),
    "ref_abmc_ilub_gmresm_d": ("i", "ppppppppp" + "pii" + "ppdiiid" + "ii"),
    "ref_bcsr_gmresm_d": ("i", "ppp" + "ppdiiid" + "ii"),
    "ref_gcrm_d": (None, "pppppdiiid"), "ref_ilu_gcrm_d": (None, "ppppppppp" + "ppdiiid"),
    "ref_ilub_gcrm_d": ("i", "ppppppppp" + "ppdiiid" + "ii"),
    "ref_bicgstab_d": (None, "pppppdiid"), "ref_ilu_bicgstab_d": (None, "ppppppppp" + "ppdiid"),
    "ref_ilub_bicgstab_d": ("i", "ppppppppp" + "ppdiid" + "ii"),
    "ref_relative_residual_error": (None, "pppppi"), "ref_flush": (None, ""),
    "ref_read_mm": ("i", "pippppii"), "ref_expand_d": ("i", "pppippppi"), "ref_remove_zeros_d": ("i", "pppi"),
}
_K = {"p": C.c_void_p, "i": C.c_int, "d": C.c_double, "f": C.c_float, None: None}


def ref():
    """The compiled reference, or None when oracle/_ref was not built."""
    global _ref, _ref_tried
    if not _ref_tried:
        _ref_tried = True
        if REF_SO.exists():
            L = C.CDLL(str(REF_SO))
            for name, (rt, args) in _REF_PROTOS.items():
                f = getattr(L, name)
                f.restype = _K[rt]
                f.argtypes = [_K[a] for a in args]
            _ref = L
    return _ref


def P(a: np.ndarray):
    """Pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class PrecSpec(C.Structure):
    _fields_ = [("prec", C.c_int), ("bnl", C.c_int), ("bnw", C.c_int),
                ("lval", C.c_void_p), ("lcind", C.c_void_p), ("lrptr", C.c_void_p),
                ("uval", C.c_void_p), ("ucind", C.c_void_p), ("urptr", C.c_void_p)]


def prec_spec(kind=0, L=None, U=None, bnl=0, bnw=0):
    """or_prec_t for the oracle's drivers; keeps the arrays alive on the object."""
    s = PrecSpec()
    s.prec, s.bnl, s.bnw = kind, bnl, bnw
    s._keep = (L, U)
    if L is not None:
        s.lval, s.lcind, s.lrptr = P(L[0]), P(L[1]), P(L[2])
        s.uval, s.ucind, s.urptr = P(U[0]), P(U[1]), P(U[2])
    return s
