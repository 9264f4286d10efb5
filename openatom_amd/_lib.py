"""ctypes binding of liboa_ortho.so (the C ABI declared in include/oa_ortho.h).

The shared library is the product: every product call goes through it.  If it
is not built, loading raises -- there is no fallback path.
"""
from __future__ import annotations

import ctypes
import os
import re

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboa_ortho.so")
HEADER = os.path.join(HERE, "..", "include", "oa_ortho.h")

OA_SUCCESS = 0
OA_ERR_HIP = -1
OA_ERR_NOT_CONVERGED = -2
OA_ERR_NO_DEVICE = -3
OA_ERR_UNSUPPORTED = -4
OA_ERR_CAPACITY = -5
OA_ERR_IO = -6

OA_PREC_FP64 = 0
OA_PREC_INT8 = 1
OA_PREC_FP16 = 2
OA_PREC_BF16 = 3
OA_PREC_FP32 = 4
OA_PREC_REUSE_DIGITS = 1 << 16
OA_PREC_LEAN_BACKMUL = 1 << 12
OA_PREC_ADAPTIVE = 1 << 13


def prec_code(prec) -> int:
    """'fp64' | 'int8x5' | 'int8x5a' | 'int8x6l' | 'fp16x3' | 'bf16x4' | int -> C
    precision code.  Suffix letters: 'l' selects the lean back-multiply (one
    digit level fewer), 'a' the adaptive per-chunk slice counts (int8, 4-6
    slices at most)."""
    if isinstance(prec, int):
        return prec
    p = prec.lower()
    if p == "fp64":
        return OA_PREC_FP64
    m = re.fullmatch(r"(int8|fp16|bf16|fp32)(?:x(\d))?(l?)(a?)", p)
    if not m:
        raise ValueError(f"unknown precision {prec!r}")
    base = {"int8": OA_PREC_INT8, "fp16": OA_PREC_FP16, "bf16": OA_PREC_BF16,
            "fp32": OA_PREC_FP32}[m.group(1)]
    slices = int(m.group(2) or 3)
    return (base | (slices << 8) | (OA_PREC_LEAN_BACKMUL if m.group(3) else 0)
            | (OA_PREC_ADAPTIVE if m.group(4) else 0))


class OaError(RuntimeError):
    def __init__(self, status: int, what: str):
        self.status = status
        super().__init__(f"{what}: {status_string(status)} (status {status})")


class OrthoNotConverged(OaError):
    """S->T hit invsqr_max_iter (the reference aborts, ortho.C:129-133)."""


class oa_ns_report(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int32), ("done", ctypes.c_int32),
                ("status", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("error", ctypes.c_double)]


class oa_instance(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("npw", ctypes.c_int), ("nkx0", ctypes.c_int),
                ("nspin", ctypes.c_int), ("psi", ctypes.c_void_p), ("ldpsi", ctypes.c_long),
                ("psi_out", ctypes.c_void_p), ("ldout", ctypes.c_long),
                ("S", ctypes.c_void_p), ("T", ctypes.c_void_p), ("tol", ctypes.c_double),
                ("maxit", ctypes.c_int), ("pad", ctypes.c_int)]


_lib = None
_vp = ctypes.c_void_p
_i = ctypes.c_int
_l = ctypes.c_long
_d = ctypes.c_double


def _sig(f, res, *args):
    f.restype = res
    f.argtypes = list(args)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built: run `python -m openatom_amd.build` "
            "(the HIP library is required; there is no CPU path)")
    L = ctypes.CDLL(LIB_PATH)
    _sig(L.oa_create, _i, ctypes.POINTER(_vp), _i)
    _sig(L.oa_destroy, _i, _vp)
    _sig(L.oa_set_stream, _i, _vp, _vp)
    _sig(L.oa_get_stream, _vp, _vp)
    _sig(L.oa_synchronize, _i, _vp)
    _sig(L.oa_reserve, _i, _vp, _i, _i)
    _sig(L.oa_reserve_prec, _i, _vp, _i, _i, _i)
    _sig(L.oa_status_string, ctypes.c_char_p, _i)
    _sig(L.oa_dgemm, _i, _vp, ctypes.c_char, ctypes.c_char, _i, _i, _i, _d, _vp, _i, _vp,
         _i, _d, _vp, _i)
    _sig(L.oa_overlap_sym, _i, _vp, _i, _i, _vp, _l, _i, _d, _i, _vp, _d, _vp)
    _sig(L.oa_slice_counts, _i, _vp, _i, _i, _vp, _l, _i, _i, _vp, _i, ctypes.POINTER(_i),
         ctypes.POINTER(_i))
    _sig(L.oa_lambda, _i, _vp, _i, _i, _vp, _l, _vp, _l, _d, _i, _vp)
    _sig(L.oa_s_to_t, _i, _vp, _i, _vp, _i, _d, _i, _vp, ctypes.POINTER(_i),
         ctypes.POINTER(_d))
    _sig(L.oa_s_to_t_async, _i, _vp, _i, _vp, _i, _d, _i, _vp, _vp)
    _sig(L.oa_backmul, _i, _vp, _i, _i, _vp, _i, _vp, _l, _d, _d, _vp, _l, _i)
    _sig(L.oa_lambda_force, _i, _vp, _i, _i, _vp, _i, _vp, _l, _i, _d, _d, _d, _vp, _l, _i)
    _sig(L.oa_scale_kx0, _i, _vp, _i, _i, _vp, _l, _i, _d)
    _sig(L.oa_zoverlap, _i, _vp, _i, _i, _vp, _l, _d, _vp, _i, _d, _vp)
    _sig(L.oa_state_read, _i, ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp, _vp,
         ctypes.POINTER(_i), ctypes.POINTER(_i), _vp)
    _sig(L.oa_state_write, _i, ctypes.c_char_p, _i, _i, _i, _vp, _vp, _vp, _vp, _vp)
    _sig(L.oa_kx0_expand, _i, _vp, _i, _i, _i, _vp, _l, _vp, _l)
    _sig(L.oa_kx0_contract, _i, _vp, _i, _i, _i, _vp, _l, _vp, _l)
    _sig(L.oa_zlambda, _i, _vp, _i, _i, _vp, _l, _vp, _l, _d, _vp, _i)
    _sig(L.oa_zs_to_t, _i, _vp, _i, _vp, _i, _d, _i, _i, _vp, ctypes.POINTER(_i),
         ctypes.POINTER(_d))
    _sig(L.oa_zbackmul, _i, _vp, _i, _i, _vp, _i, _vp, _l, _d, _d, _vp, _l, _i)
    _sig(L.oa_zortho_step, _i, _vp, _i, _i, _vp, _l, _i, _d, _i, _i, _vp, _vp, _vp, _l,
         ctypes.POINTER(_i), ctypes.POINTER(_d), _i)
    _sig(L.oa_ortho_step, _i, _vp, _i, _i, _vp, _l, _i, _i, _d, _i, _i, _vp, _vp, _vp, _l,
         ctypes.POINTER(_i), ctypes.POINTER(_d))
    _sig(L.oa_ortho_step_noalloc, _i, _vp, _i, _i, _vp, _l, _i, _i, _d, _i, _i, _vp, _vp, _vp,
         _l, ctypes.POINTER(_i), ctypes.POINTER(_d))
    _sig(L.oa_noalloc_scratch_bytes, ctypes.c_size_t)
    _dp = ctypes.POINTER(ctypes.c_double)
    _sig(L.oa_zgemm, _i, _vp, ctypes.c_char, ctypes.c_char, _i, _i, _i, _dp, _vp, _i, _vp, _i,
         _dp, _vp, _i)
    _pi, _pl = ctypes.POINTER(_i), ctypes.POINTER(_l)
    _sig(L.oa_collate_rows, _i, _vp, _i, _pi, _pi, _pl, _pi, ctypes.POINTER(_vp), _i, _vp, _l,
         _i)
    _sig(L.oa_scatter_rows, _i, _vp, _i, _pi, _pi, _pl, _pi, ctypes.POINTER(_vp), _i, _pi,
         _vp, _l, _i)
    _sig(L.oa_batch_create, _i, _vp, ctypes.POINTER(_vp))
    _sig(L.oa_batch_destroy, _i, _vp)
    _sig(L.oa_batch_plan, _i, _vp, _i, ctypes.POINTER(oa_instance), _i)
    _sig(L.oa_batch_footprint, ctypes.c_size_t, _vp)
    _sig(L.oa_batch_reserve, _i, _vp, ctypes.c_size_t, ctypes.c_size_t)
    _sig(L.oa_instance_footprint, ctypes.c_size_t, _i, _i, _i)
    _sig(L.oa_batch_run, _i, _vp)
    _sig(L.oa_batch_launch, _i, _vp)
    _sig(L.oa_batch_reports, _i, _vp, ctypes.POINTER(oa_ns_report), _i)
    _sig(L.oa_batch_tune, _i, _vp, _i)
    _sig(L.oa_batch_config, _i, _vp, _i, ctypes.POINTER(_i), ctypes.POINTER(_i))
    _sig(L.oa_batch_get_table, _i, _vp, ctypes.c_char_p, _i)
    _sig(L.oa_batch_set_table, _i, _vp, ctypes.c_char_p)
    _lib = L
    return L


def status_string(st: int) -> str:
    try:
        return lib().oa_status_string(st).decode()
    except ImportError:
        return str(st)


def check(st: int, what: str) -> None:
    if st == OA_SUCCESS:
        return
    if st == OA_ERR_NOT_CONVERGED:
        raise OrthoNotConverged(st, what)
    if st > 0:
        raise ValueError(f"{what}: invalid argument #{st}")
    raise OaError(st, what)


def declared_symbols() -> list[str]:
    """Every function the public header declares (for the export test)."""
    txt = open(HEADER).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[\w\s\*]+?\b(\w+)\s*\(", txt, flags=re.M)
    return sorted({n for n in names if n not in ("if", "defined")})
