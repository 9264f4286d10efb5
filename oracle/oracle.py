"""Python face of the CPU oracle (test infrastructure only).

Loads ``oracle/liboa_oracle.so`` (the C restatement in ``oa_oracle.c``) and,
when it was built, ``oracle/_ref/librefblas.so`` (the reference's own netlib
BLAS compiled from ``src_mathlib/*.f``).  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import
this module; the product package never does.

All arrays are numpy float64, C-contiguous.  States are row-major ``[n][2*npw]``
(complex coefficients interleaved re, im), the layout the reference hands to
DGEMM (MessageDataCollator.h:222; ckPairCalculator.C:871-872).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboa_oracle.so")
REF_SO = os.path.join(HERE, "_ref", "librefblas.so")

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_cp = ctypes.c_char_p

DGEMM_FN = ctypes.CFUNCTYPE(None, _cp, _cp, _ip, _ip, _ip, _dp, _dp, _ip, _dp,
                            _ip, _dp, _dp, _ip)


def build(force: bool = False) -> None:
    """Compile the oracle (and the reference BLAS when the reference tree is
    present) with the committed Makefile."""
    if force or not os.path.exists(ORACLE_SO):
        subprocess.run(["make", "-s", "-f", os.path.join(HERE, "Makefile"),
                        "all"], check=True)


_lib = None
_ref = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(ORACLE_SO):
            build()
        L = ctypes.CDLL(ORACLE_SO)
        L.oao_dgemm_info.restype = ctypes.c_int
        L.oao_dgemm_info.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                     _dp, ctypes.c_int, _dp, ctypes.c_int,
                                     ctypes.c_double, _dp, ctypes.c_int]
        L.oao_overlap.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                  _dp, ctypes.c_int, ctypes.c_void_p]
        L.oao_lambda.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp, _dp,
                                 ctypes.c_void_p]
        L.oao_s_to_t.restype = ctypes.c_int
        L.oao_s_to_t.argtypes = [ctypes.c_int, _dp, ctypes.c_int, ctypes.c_double,
                                 ctypes.c_int, _dp, _ip, _dp, ctypes.c_void_p]
        L.oao_piny_iter_tmat.restype = ctypes.c_int
        L.oao_piny_iter_tmat.argtypes = [_dp, _dp, ctypes.c_int, _ip, _dp,
                                         ctypes.c_void_p]
        L.oao_backmul.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                  _dp, ctypes.c_double, ctypes.c_double, _dp,
                                  ctypes.c_int, ctypes.c_void_p]
        L.oao_ortho_step.restype = ctypes.c_int
        L.oao_ortho_step.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                     _dp, _dp, _dp, _ip, _dp, ctypes.c_int,
                                     ctypes.c_void_p]
        L.oao_orthonormality_residual.restype = ctypes.c_double
        L.oao_orthonormality_residual.argtypes = [ctypes.c_int, ctypes.c_int, _dp,
                                                  ctypes.c_int, ctypes.c_double,
                                                  ctypes.c_int, ctypes.c_void_p]
        L.oao_cla_multiply.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                       ctypes.c_double, _dp, _dp, ctypes.c_double,
                                       _dp, ctypes.c_void_p]
        L.oao_max_threads.restype = ctypes.c_int
        L.oao_set_threads.argtypes = [ctypes.c_int]
        L.oao_zgemm_info.argtypes = [ctypes.c_char, ctypes.c_char, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, _dp, _dp, ctypes.c_int, _dp,
                                     ctypes.c_int, _dp, _dp, ctypes.c_int]
        L.oao_zgemm_info.restype = ctypes.c_int
        _vp = ctypes.c_void_p
        _i, _d = ctypes.c_int, ctypes.c_double
        L.oao_zoverlap.argtypes = [_i, _i, _dp, _dp, _vp]
        L.oao_zlambda.argtypes = [_i, _i, _dp, _dp, _dp, _vp]
        L.oao_zs_to_t.restype = _i
        L.oao_zs_to_t.argtypes = [_i, _dp, _i, _d, _i, _i, _dp, _ip, _dp, _vp]
        L.oao_zbackmul.argtypes = [_i, _i, _dp, _i, _dp, _d, _d, _dp, _vp]
        L.oao_zortho_step.restype = _i
        L.oao_zortho_step.argtypes = [_i, _i, _dp, _i, _d, _i, _i, _dp, _dp, _dp, _ip, _dp,
                                      _vp]
        L.oao_gsp_unpack.argtypes = [ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int,
                                     ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
        L.oao_array_diag_max.restype = _d
        L.oao_array_diag_max.argtypes = [_i, _dp, _i, _i, _i]
        _lib = L
    return _lib


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref() -> ctypes.CDLL:
    """The reference's compiled netlib BLAS (dgemm_, zgemm_, genmatmul_)."""
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(
                f"{REF_SO} is not built (the reference tree was absent when "
                "oracle/Makefile ran)")
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def gemm_ptr(which: str):
    """A C function pointer usable wherever the oracle takes a dgemm:
    'oracle' -> oao_dgemm, 'ref' -> the reference dgemm_."""
    if which == "oracle":
        return ctypes.cast(lib().oao_dgemm, ctypes.c_void_p)
    if which == "ref":
        return ctypes.cast(ref().dgemm_, ctypes.c_void_p)
    raise ValueError(which)


def _p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def dgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, which="oracle"):
    """Column-major DGEMM on flat buffers through either implementation.
    Returns INFO for the oracle (the reference's XERBLA is silent)."""
    if which == "oracle":
        return lib().oao_dgemm_info(ta.encode(), tb.encode(), m, n, k, alpha,
                                    _p(A), lda, _p(B), ldb, beta, _p(C), ldc)
    f = ref().dgemm_
    f.restype = None
    ci = ctypes.c_int
    cd = ctypes.c_double
    f(ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()),
      ctypes.byref(ci(m)), ctypes.byref(ci(n)), ctypes.byref(ci(k)),
      ctypes.byref(cd(alpha)), _p(A), ctypes.byref(ci(lda)), _p(B),
      ctypes.byref(ci(ldb)), ctypes.byref(cd(beta)), _p(C), ctypes.byref(ci(ldc)))
    return 0


def _pz(a: np.ndarray):
    assert a.dtype == np.complex128 and a.flags.c_contiguous
    return a.ctypes.data_as(_dp)


def zgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc, which="oracle"):
    """Column-major ZGEMM on flat complex128 buffers through either the
    restatement or the reference's compiled zgemm_ (zgemm.f)."""
    al = np.array([complex(alpha)], dtype=np.complex128)
    be = np.array([complex(beta)], dtype=np.complex128)
    if which == "oracle":
        return lib().oao_zgemm_info(ta.encode(), tb.encode(), m, n, k, _pz(al), _pz(A), lda,
                                    _pz(B), ldb, _pz(be), _pz(C), ldc)
    f = ref().zgemm_
    f.restype = None
    ci = ctypes.c_int
    f(ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()),
      ctypes.byref(ci(m)), ctypes.byref(ci(n)), ctypes.byref(ci(k)),
      _pz(al), _pz(A), ctypes.byref(ci(lda)), _pz(B), ctypes.byref(ci(ldb)), _pz(be), _pz(C),
      ctypes.byref(ci(ldc)))
    return 0


def overlap(psi: np.ndarray, nkx0: int = 1, nplanes: int = 1, gemm="oracle") -> np.ndarray:
    n, K = psi.shape
    S = np.empty((n, n))
    lib().oao_overlap(n, K // 2, _p(psi), nkx0, _p(S), nplanes, gemm_ptr(gemm))
    return S


def lambda_fwd(psi: np.ndarray, F: np.ndarray, gemm="oracle") -> np.ndarray:
    n, K = psi.shape
    L = np.empty((n, n))
    lib().oao_lambda(n, K // 2, _p(psi), _p(F), _p(L), gemm_ptr(gemm))
    return L


def s_to_t(S: np.ndarray, nspin: int = 1, tol: float = 1e-15, maxit: int = 1000,
           gemm="oracle"):
    """Returns (T, iterations, err, status); status 1 = iteration limit."""
    n = S.shape[0]
    S = np.ascontiguousarray(S)
    T = np.empty((n, n))
    it = ctypes.c_int(0)
    err = ctypes.c_double(0)
    st = lib().oao_s_to_t(n, _p(S), nspin, tol, maxit, _p(T), ctypes.byref(it),
                          ctypes.byref(err), gemm_ptr(gemm))
    return T, it.value, err.value, st


def piny_iter_tmat(S: np.ndarray, gemm="oracle"):
    n = S.shape[0]
    S = np.ascontiguousarray(S)
    T = np.empty((n, n))
    it = ctypes.c_int(0)
    tol = ctypes.c_double(0)
    st = lib().oao_piny_iter_tmat(_p(S), _p(T), n, ctypes.byref(it),
                                  ctypes.byref(tol), gemm_ptr(gemm))
    return T, it.value, tol.value, st


def backmul(T: np.ndarray, psi: np.ndarray, transT: bool = False, alpha: float = 1.0,
            beta: float = 0.0, out: np.ndarray | None = None, nplanes: int = 1,
            gemm="oracle") -> np.ndarray:
    n, K = psi.shape
    T = np.ascontiguousarray(T)
    if out is None:
        out = np.zeros((n, K))
    lib().oao_backmul(n, K, _p(T), 1 if transT else 0, _p(psi), alpha, beta,
                      _p(out), nplanes, gemm_ptr(gemm))
    return out


def ortho_step(psi: np.ndarray, nkx0: int = 1, nspin: int = 1, tol: float = 1e-15,
               maxit: int = 1000, nplanes: int = 1, gemm="oracle"):
    """One reference symmetric step.  Returns dict(S, T, psi_out, iters, err, status)."""
    n, K = psi.shape
    S = np.empty((n, n))
    T = np.empty((n, n))
    out = np.empty((n, K))
    it = ctypes.c_int(0)
    err = ctypes.c_double(0)
    st = lib().oao_ortho_step(n, K // 2, _p(psi), nkx0, nspin, tol, maxit, _p(S),
                              _p(T), _p(out), ctypes.byref(it), ctypes.byref(err),
                              nplanes, gemm_ptr(gemm))
    return dict(S=S, T=T, psi_out=out, iters=it.value, err=err.value, status=st)


def zgemm_ptr(which: str):
    """'oracle' -> oao_zgemm, 'ref' -> the reference's compiled zgemm_."""
    if which == "oracle":
        return ctypes.cast(lib().oao_zgemm, ctypes.c_void_p)
    if which == "ref":
        return ctypes.cast(ref().zgemm_, ctypes.c_void_p)
    raise ValueError(which)


def _zsq(M: np.ndarray) -> np.ndarray:
    M = np.ascontiguousarray(M, dtype=np.complex128)
    return M


def _pc(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def zoverlap(psi: np.ndarray, gemm="oracle") -> np.ndarray:
    """Complex-math forward S = psi psi^H (n x n complex128); psi is the
    interleaved [n][2*npw] float64 block."""
    n, K = psi.shape
    psi = np.ascontiguousarray(psi)
    S = np.empty((n, n), dtype=np.complex128)
    lib().oao_zoverlap(n, K // 2, _p(psi), _pc(S), zgemm_ptr(gemm))
    return S


def zlambda(psi: np.ndarray, F: np.ndarray, gemm="oracle") -> np.ndarray:
    n, K = psi.shape
    psi, F = np.ascontiguousarray(psi), np.ascontiguousarray(F)
    L = np.empty((n, n), dtype=np.complex128)
    lib().oao_zlambda(n, K // 2, _p(psi), _p(F), _pc(L), zgemm_ptr(gemm))
    return L


def zs_to_t(S: np.ndarray, nspin: int = 1, tol: float = 1e-15, maxit: int = 1000,
            modulus: bool = False, gemm="oracle"):
    """Complex Ortho S->T.  Returns (T, iterations, error, status)."""
    S = _zsq(S)
    n = S.shape[0]
    T = np.empty_like(S)
    it = ctypes.c_int(0)
    err = ctypes.c_double(0)
    st = lib().oao_zs_to_t(n, _pc(S), nspin, tol, maxit, 1 if modulus else 0, _pc(T),
                           ctypes.byref(it), ctypes.byref(err), zgemm_ptr(gemm))
    return T, it.value, err.value, st


def zbackmul(M: np.ndarray, psi: np.ndarray, transM: bool = False, alpha: float = 1.0,
             beta: float = 0.0, out: np.ndarray | None = None, gemm="oracle") -> np.ndarray:
    n, K = psi.shape
    M = _zsq(M)
    psi = np.ascontiguousarray(psi)
    if out is None:
        out = np.zeros((n, K))
    lib().oao_zbackmul(n, K // 2, _pc(M), 1 if transM else 0, _p(psi), alpha, beta, _p(out),
                       zgemm_ptr(gemm))
    return out


def zortho_step(psi: np.ndarray, nspin: int = 1, tol: float = 1e-15, maxit: int = 1000,
                modulus: bool = False, gemm="oracle"):
    n, K = psi.shape
    psi = np.ascontiguousarray(psi)
    S = np.empty((n, n), dtype=np.complex128)
    T = np.empty_like(S)
    out = np.empty((n, K))
    it = ctypes.c_int(0)
    err = ctypes.c_double(0)
    st = lib().oao_zortho_step(n, K // 2, _p(psi), nspin, tol, maxit, 1 if modulus else 0,
                               _pc(S), _pc(T), _p(out),
                               ctypes.byref(it), ctypes.byref(err), zgemm_ptr(gemm))
    return dict(S=S, T=T, psi_out=out, iters=it.value, err=err.value, status=st)


def gsp_unpack(data: np.ndarray, force: np.ndarray, nkx0: int = 1, double_pack: bool = True,
               mode: int = 0, lsda: bool = False) -> np.ndarray:
    """GSpace accumulation of a backward result into the force, in place
    (mode 0 partialResultMsg, 1 CkReductionMsg, 2 dynamics)."""
    n, K = data.shape
    assert force.shape == data.shape and force.flags.c_contiguous
    data = np.ascontiguousarray(data)
    lib().oao_gsp_unpack(n, K // 2, _p(data), nkx0, 1 if double_pack else 0, mode,
                         1 if lsda else 0, _p(force))
    return force


def array_diag_max(S: np.ndarray, lsda: bool = False, og: int = 0) -> float:
    """Ortho::array_diag_max over og x og Ortho tiles (og = 0: one tile),
    reduced with max_double as maxCheck receives it; complex S takes the
    modulus (myabs = std::abs)."""
    n = S.shape[0]
    cplx = np.iscomplexobj(S)
    A = np.ascontiguousarray(S.view(np.float64) if cplx else S, dtype=np.float64)
    return float(lib().oao_array_diag_max(n, _p(A), og, 1 if lsda else 0, 1 if cplx else 0))


def orthonormality_residual(psi: np.ndarray, nkx0: int = 1, occ: float = 1.0,
                            nplanes: int = 1, gemm="oracle") -> float:
    n, K = psi.shape
    return lib().oao_orthonormality_residual(n, K // 2, _p(psi), nkx0, occ, nplanes,
                                             gemm_ptr(gemm))


def set_threads(t: int) -> None:
    lib().oao_set_threads(t)


def max_threads() -> int:
    return lib().oao_max_threads()
