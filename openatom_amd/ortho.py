"""Host-side mirror of the reference's PairCalculator / Ortho interface for the
orthonormalization path, over the C ABI of liboa_ortho.so.

Device memory comes from PyTorch tensors (float64, CUDA/HIP device); the
library runs every product on the GPU.  Names follow the reference:

  PairCalculator.multiplyForward   ckPairCalculator.C:821-933 (symmetric
                                   S with alpha=2 and the kx=0 weight, or the
                                   asymmetric Lambda with alpha=1)
  PairCalculator.multiplyResult    ckPairCalculator.C:1201-1468 ->
                                   bwMultiplyHelper :1710-1946
  PairCalculator.bwMultiplyDynOrthoT  ckPairCalculator.C:1965-2031 (-F*T)
  PairCalculator.multiplyPsiV      ckPairCalculator.C:1128-1176
  Ortho.start_calc                 ortho.C:153-266 + do_iteration..collect_error
                                   (:848-971, :109-146)
  Ortho.diag                       ortho.h:282-313 (the diagonal
                                   array_diag_max measures against)
  Ortho.acceptSectionLambda        ortho.C:497-608 (lsda: Lambda x 2)
  Ortho.gamma                      ortho.C:565-570 (Gamma = T * Lambda)
"""
from __future__ import annotations

import ctypes
import math
import weakref

import torch

from . import _lib
from ._lib import check, prec_code


def _ptr(t: torch.Tensor | None) -> int | None:
    return None if t is None else t.data_ptr()


def _ld(t: torch.Tensor) -> int:
    """Row pitch in elements (torch reports arbitrary strides for 1-row views)."""
    return t.stride(0) if t.shape[0] > 1 else t.shape[1]


def _need(t: torch.Tensor, name: str, rows: int | None = None) -> None:
    if not t.is_cuda or t.dtype != torch.float64:
        raise TypeError(f"{name} must be a float64 device tensor")
    if t.dim() != 2 or (t.stride(1) != 1 and t.shape[1] > 1):
        raise ValueError(f"{name} must be 2-D with unit column stride")
    if rows is not None and t.shape[0] != rows:
        raise ValueError(f"{name} must have {rows} rows")


class Handle:
    """An oa_handle_t bound to a device and (by default) to PyTorch's current
    stream on that device, so library work orders with torch work."""

    def __init__(self, device: int | None = None, stream: torch.cuda.Stream | None = None):
        L = _lib.lib()
        self.device = torch.cuda.current_device() if device is None else device
        h = ctypes.c_void_p()
        check(L.oa_create(ctypes.byref(h), self.device), "oa_create")
        self.h = h
        # batches made on this handle: destroyed before it (a batch's
        # teardown synchronizes on the handle's device and stream)
        self._batches = weakref.WeakSet()
        self.use_stream(stream)

    def use_stream(self, stream: torch.cuda.Stream | None = None) -> None:
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        self.stream = s
        check(_lib.lib().oa_set_stream(self.h, ctypes.c_void_p(s.cuda_stream)), "oa_set_stream")

    def reserve(self, n: int, npw: int, prec="fp64") -> None:
        check(_lib.lib().oa_reserve_prec(self.h, n, npw, prec_code(prec)), "oa_reserve_prec")

    def synchronize(self) -> None:
        check(_lib.lib().oa_synchronize(self.h), "oa_synchronize")

    def close(self) -> None:
        if getattr(self, "h", None):
            for b in list(getattr(self, "_batches", ())):
                b.close()
            _lib.lib().oa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -------------------------------------------------------------- raw calls
    def dgemm(self, transa: str, transb: str, m: int, n: int, k: int, alpha: float,
              a: torch.Tensor, lda: int, b: torch.Tensor, ldb: int, beta: float,
              c: torch.Tensor, ldc: int) -> int:
        """Column-major DGEMM on device buffers (returns the BLAS INFO)."""
        return _lib.lib().oa_dgemm(self.h, transa.encode(), transb.encode(), m, n, k, alpha,
                                   _ptr(a), lda, _ptr(b), ldb, beta, _ptr(c), ldc)

    def zgemm(self, transa: str, transb: str, m: int, n: int, k: int, alpha: complex,
              a: torch.Tensor, lda: int, b: torch.Tensor, ldb: int, beta: complex,
              c: torch.Tensor, ldc: int) -> int:
        """Column-major ZGEMM on complex128 device buffers (returns the BLAS
        INFO); the complex-math PairCalculator's GEMM."""
        al = (ctypes.c_double * 2)(complex(alpha).real, complex(alpha).imag)
        be = (ctypes.c_double * 2)(complex(beta).real, complex(beta).imag)
        return _lib.lib().oa_zgemm(self.h, transa.encode(), transb.encode(), m, n, k, al,
                                   _ptr(a), lda, _ptr(b), ldb, be, _ptr(c), ldc)

    def overlap_sym(self, psi: torch.Tensor, nkx0: int = 1, alpha: float = 2.0,
                    prec="fp64", S: torch.Tensor | None = None, diag: float = 0.0,
                    maxdev: torch.Tensor | None = None) -> torch.Tensor:
        n = psi.shape[0]
        _need(psi, "psi")
        npw = psi.shape[1] // 2
        if S is None:
            S = torch.empty((n, n), dtype=torch.float64, device=psi.device)
        check(_lib.lib().oa_overlap_sym(self.h, n, npw, _ptr(psi), _ld(psi), nkx0, alpha,
                                        prec_code(prec), _ptr(S), diag, _ptr(maxdev)),
              "oa_overlap_sym")
        return S

    def slice_counts(self, psi: torch.Tensor, nkx0: int = 1, prec="int8x5a"):
        """Per chunk digit counts an adaptive split precision assigns to psi.
        Returns (counts as a list, columns per chunk in doubles)."""
        n = psi.shape[0]
        _need(psi, "psi")
        npw = psi.shape[1] // 2
        nz, kc = ctypes.c_int(0), ctypes.c_int(0)
        L = _lib.lib()
        st = L.oa_slice_counts(self.h, n, npw, _ptr(psi), _ld(psi), nkx0, prec_code(prec),
                               None, 0, ctypes.byref(nz), ctypes.byref(kc))
        if st != _lib.OA_ERR_CAPACITY:
            check(st, "oa_slice_counts")
        buf = (ctypes.c_int * max(nz.value, 1))()
        check(L.oa_slice_counts(self.h, n, npw, _ptr(psi), _ld(psi), nkx0, prec_code(prec),
                                buf, nz.value, ctypes.byref(nz), ctypes.byref(kc)),
              "oa_slice_counts")
        return list(buf[:nz.value]), kc.value

    def lambda_fwd(self, psi: torch.Tensor, F: torch.Tensor, alpha: float = 1.0,
                   prec="fp64", L: torch.Tensor | None = None) -> torch.Tensor:
        n = psi.shape[0]
        _need(psi, "psi")
        _need(F, "F", n)
        npw = psi.shape[1] // 2
        if L is None:
            L = torch.empty((n, n), dtype=torch.float64, device=psi.device)
        check(_lib.lib().oa_lambda(self.h, n, npw, _ptr(psi), _ld(psi), _ptr(F),
                                   _ld(F), alpha, prec_code(prec), _ptr(L)), "oa_lambda")
        return L

    def s_to_t(self, S: torch.Tensor, nspin: int = 1, tol: float = 1e-15, maxit: int = 1000,
               T: torch.Tensor | None = None):
        """Synchronous S->T.  Returns (T, iterations, error); raises
        OrthoNotConverged when the iteration limit was reached."""
        n = S.shape[0]
        _need(S, "S", n)
        if T is None:
            T = torch.empty_like(S)
        it = ctypes.c_int(0)
        err = ctypes.c_double(0)
        st = _lib.lib().oa_s_to_t(self.h, n, _ptr(S), nspin, tol, maxit, _ptr(T),
                                  ctypes.byref(it), ctypes.byref(err))
        self.last_iterations, self.last_error = it.value, err.value
        check(st, "oa_s_to_t")
        return T, it.value, err.value

    def s_to_t_async(self, S: torch.Tensor, T: torch.Tensor, report: torch.Tensor,
                     nspin: int = 1, tol: float = 1e-15, maxit: int = 12) -> None:
        """Enqueue `maxit` guarded iterations; `report` is a 24-byte device
        tensor receiving an oa_ns_report."""
        n = S.shape[0]
        check(_lib.lib().oa_s_to_t_async(self.h, n, _ptr(S), nspin, tol, maxit, _ptr(T),
                                         _ptr(report)), "oa_s_to_t_async")

    def backmul(self, M: torch.Tensor, psi: torch.Tensor, transM: bool = False,
                alpha: float = 1.0, beta: float = 0.0, out: torch.Tensor | None = None,
                prec="fp64", reuse_digits: bool = False) -> torch.Tensor:
        """reuse_digits: take psi's digit slices from the preceding split
        slicing of the same psi on this handle (split precisions only); psi is
        sliced again when the workspace no longer holds them."""
        n = psi.shape[0]
        _need(psi, "psi")
        _need(M, "M", n)
        npw = psi.shape[1] // 2
        if out is None:
            out = torch.zeros_like(psi) if beta != 0.0 else torch.empty_like(psi)
        pc = prec_code(prec)
        if reuse_digits and (pc & 0xff) != _lib.OA_PREC_FP64:
            pc |= _lib.OA_PREC_REUSE_DIGITS
        check(_lib.lib().oa_backmul(self.h, n, npw, _ptr(M), 1 if transM else 0, _ptr(psi),
                                    _ld(psi), alpha, beta, _ptr(out), _ld(out), pc),
              "oa_backmul")
        return out

    def lambda_force(self, M: torch.Tensor, psi: torch.Tensor, force: torch.Tensor,
                     transM: bool = False, nkx0: int = 1, wkx0: float = 0.5,
                     alpha: float = -1.0, beta: float = 1.0, prec="fp64") -> torch.Tensor:
        """force := beta force + alpha w .* op(M) psi in place, w = wkx0 on the
        leading 2*nkx0 doubles of each row and 1 elsewhere."""
        n = psi.shape[0]
        _need(psi, "psi")
        _need(M, "M", n)
        _need(force, "force", n)
        npw = psi.shape[1] // 2
        check(_lib.lib().oa_lambda_force(self.h, n, npw, _ptr(M), 1 if transM else 0, _ptr(psi),
                                         _ld(psi), nkx0, wkx0, alpha, beta, _ptr(force),
                                         _ld(force), prec_code(prec)), "oa_lambda_force")
        return force

    # ---- complex-math (k-point) path: states [n][2*npw] interleaved complex,
    # n x n matrices as complex128 tensors (or [n][2n] float64)

    @staticmethod
    def _zmat(M: torch.Tensor, n: int, name: str) -> torch.Tensor:
        if M.dtype == torch.complex128:
            M = torch.view_as_real(M).reshape(n, 2 * n)
        _need(M, name, n)
        if M.shape[1] != 2 * n or not M.is_contiguous():
            raise ValueError(f"{name} must be a contiguous n x n complex matrix")
        return M

    def zoverlap(self, psi: torch.Tensor, alpha: float = 1.0, prec="fp64",
                 S: torch.Tensor | None = None, diag: float = 0.0,
                 maxdev: torch.Tensor | None = None) -> torch.Tensor:
        """S = alpha psi psi^H as an n x n complex128 tensor; `maxdev` (a
        1-element float64 device tensor) receives array_diag_max of S."""
        _need(psi, "psi")
        n = psi.shape[0]
        if S is None:
            S = torch.empty((n, n), dtype=torch.complex128, device=psi.device)
        check(_lib.lib().oa_zoverlap(self.h, n, psi.shape[1] // 2, _ptr(psi), _ld(psi), alpha,
                                     S.data_ptr(), prec_code(prec), diag, _ptr(maxdev)),
              "oa_zoverlap")
        return S

    def zlambda(self, psi: torch.Tensor, F: torch.Tensor, alpha: float = 1.0,
                prec="fp64") -> torch.Tensor:
        """L = alpha psi F^H as an n x n complex128 tensor."""
        _need(psi, "psi")
        n = psi.shape[0]
        _need(F, "F", n)
        L = torch.empty((n, n), dtype=torch.complex128, device=psi.device)
        check(_lib.lib().oa_zlambda(self.h, n, psi.shape[1] // 2, _ptr(psi), _ld(psi), _ptr(F),
                                    _ld(F), alpha, L.data_ptr(), prec_code(prec)), "oa_zlambda")
        return L

    def zs_to_t(self, S: torch.Tensor, nspin: int = 1, tol: float = 1e-15, maxit: int = 1000,
                modulus: bool = False, raise_on_limit: bool = True):
        """Complex Ortho S->T.  Returns (T complex128, iterations, error)."""
        n = S.shape[0]
        Sv = self._zmat(S, n, "S")
        T = torch.empty((n, n), dtype=torch.complex128, device=S.device)
        it = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        st = _lib.lib().oa_zs_to_t(self.h, n, _ptr(Sv), nspin, tol, maxit, 1 if modulus else 0,
                                   T.data_ptr(), ctypes.byref(it), ctypes.byref(err))
        self.last_iterations, self.last_error = it.value, err.value
        if raise_on_limit or st != _lib.OA_ERR_NOT_CONVERGED:
            check(st, "oa_zs_to_t")
        return T, it.value, err.value

    def zbackmul(self, M: torch.Tensor, psi: torch.Tensor, transM: bool = False,
                 alpha: float = 1.0, beta: float = 0.0,
                 out: torch.Tensor | None = None, prec="fp64") -> torch.Tensor:
        """out = alpha op(M) psi + beta out, op(M) = M or M^H."""
        _need(psi, "psi")
        n = psi.shape[0]
        Mv = self._zmat(M, n, "M")
        if out is None:
            out = torch.zeros_like(psi) if beta != 0.0 else torch.empty_like(psi)
        check(_lib.lib().oa_zbackmul(self.h, n, psi.shape[1] // 2, _ptr(Mv), 1 if transM else 0,
                                     _ptr(psi), _ld(psi), alpha, beta, _ptr(out), _ld(out),
                                     prec_code(prec)), "oa_zbackmul")
        return out

    def zortho_step(self, psi: torch.Tensor, nspin: int = 1, tol: float = 1e-15,
                    maxit: int = 1000, modulus: bool = False, prec="fp64"):
        """One complex step.  Returns (S, T, psi_out, iterations, error).
        modulus: stop on sum |dA|^2 instead of the reference's Re sum dA^2."""
        _need(psi, "psi")
        n = psi.shape[0]
        S = torch.empty((n, n), dtype=torch.complex128, device=psi.device)
        T = torch.empty_like(S)
        out = torch.empty_like(psi)
        it = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        st = _lib.lib().oa_zortho_step(self.h, n, psi.shape[1] // 2, _ptr(psi), _ld(psi), nspin,
                                       tol, maxit, 1 if modulus else 0, S.data_ptr(),
                                       T.data_ptr(), _ptr(out),
                                       _ld(out), ctypes.byref(it), ctypes.byref(err),
                                       prec_code(prec))
        self.last_iterations, self.last_error = it.value, err.value
        check(st, "oa_zortho_step")
        return S, T, out, it.value, err.value

    def scale_kx0(self, X: torch.Tensor, nkx0: int, s: float) -> torch.Tensor:
        """X[:, :2*nkx0] *= s in place."""
        _need(X, "X")
        check(_lib.lib().oa_scale_kx0(self.h, X.shape[0], X.shape[1] // 2, _ptr(X), _ld(X),
                                      nkx0, s), "oa_scale_kx0")
        return X

    def ortho_step(self, psi: torch.Tensor, nkx0: int = 1, nspin: int = 1,
                   tol: float = 1e-15, maxit: int = 1000, prec="fp64",
                   S=None, T=None, out=None, sync: bool = True):
        """One symmetric step.  Returns (S, T, psi_out, iterations, error);
        iterations / error are None when sync is False."""
        n = psi.shape[0]
        _need(psi, "psi")
        npw = psi.shape[1] // 2
        dev = psi.device
        S = torch.empty((n, n), dtype=torch.float64, device=dev) if S is None else S
        T = torch.empty((n, n), dtype=torch.float64, device=dev) if T is None else T
        out = torch.empty_like(psi) if out is None else out
        it = ctypes.c_int(0)
        err = ctypes.c_double(0)
        st = _lib.lib().oa_ortho_step(self.h, n, npw, _ptr(psi), _ld(psi), nkx0, nspin,
                                      tol, maxit, prec_code(prec), _ptr(S), _ptr(T), _ptr(out),
                                      _ld(out), ctypes.byref(it) if sync else None,
                                      ctypes.byref(err) if sync else None)
        check(st, "oa_ortho_step")
        if sync:
            return S, T, out, it.value, err.value
        return S, T, out, None, None

    # ------------------------------------------------------------ transport
    @staticmethod
    def _msg_table(msgs, block=None):
        for m in msgs:
            r0, nr, c0, nc, buf = m[:5]
            if buf.numel() != nr * nc:
                raise ValueError(f"message buffer holds {buf.numel()} doubles, not {nr} x {nc}")
            if buf.dtype != torch.float64:
                raise TypeError("message buffers must be float64")
            if block is not None and (c0 < 0 or c0 + nc > block.shape[1]):
                raise ValueError("invalid argument #5: message columns outside the block")
        count = len(msgs)
        ci, cl = ctypes.c_int * max(count, 1), ctypes.c_long * max(count, 1)
        return (count, ci(*[m[0] for m in msgs]), ci(*[m[1] for m in msgs]),
                cl(*[m[2] for m in msgs]), ci(*[m[3] for m in msgs]))

    def collate_rows(self, block: torch.Tensor, msgs) -> torch.Tensor:
        """MessageDataCollator::operator() (MessageDataCollator.h:163-247) for a
        list of messages (row0, nrows, col0, ncols, data): each data tensor
        (nrows x ncols float64, host or device, all on the same side) is
        copied into block[row0:row0+nrows, col0:col0+ncols]."""
        _need(block, "block")
        count, r0, nr, c0, nc = self._msg_table(msgs, block)
        datas = [m[4].contiguous() for m in msgs]
        on_host = bool(datas) and not datas[0].is_cuda
        if any(d.is_cuda == on_host for d in datas):
            raise ValueError("message buffers must all be host or all be device tensors")
        ptrs = (ctypes.c_void_p * max(count, 1))(*[d.data_ptr() for d in datas])
        check(_lib.lib().oa_collate_rows(self.h, count, r0, nr, c0, nc, ptrs, int(on_host),
                                         _ptr(block), _ld(block), block.shape[0]),
              "oa_collate_rows")
        return block

    def scatter_rows(self, block: torch.Tensor, msgs) -> None:
        """The return leg of the backward path (sendBWResultDirect,
        ckPairCalculator.C:2365-2452, unpacked by CP_State_GSpacePlane.C:
        3120-3126 / 3197-3203): msgs are (row0, nrows, col0, ncols, dst,
        accumulate); dst (nrows x ncols float64, host or device) receives
        block[row0:row0+nrows, col0:col0+ncols], copied or added (fastAdd)."""
        _need(block, "block")
        count, r0, nr, c0, nc = self._msg_table(msgs, block)
        dsts = [m[4] for m in msgs]
        for d in dsts:
            if not d.is_contiguous():
                raise ValueError("destination buffers must be contiguous")
        on_host = bool(dsts) and not dsts[0].is_cuda
        if any(d.is_cuda == on_host for d in dsts):
            raise ValueError("destinations must all be host or all be device tensors")
        ptrs = (ctypes.c_void_p * max(count, 1))(*[d.data_ptr() for d in dsts])
        acc = (ctypes.c_int * max(count, 1))(*[1 if m[5] else 0 for m in msgs])
        check(_lib.lib().oa_scatter_rows(self.h, count, r0, nr, c0, nc, ptrs, int(on_host), acc,
                                         _ptr(block), _ld(block), block.shape[0]),
              "oa_scatter_rows")


class Batch:
    """Many independent instances (beads, k-points, tempering replicas) of
    uneven size advanced together: each phase of the step is one persistent
    grouped launch over all of them (oa_batch_*).  Each instance's results
    equal Handle.ortho_step on that instance alone, bit for bit."""

    def __init__(self, handle: Handle | None = None):
        self.handle = handle or default_handle()
        if not getattr(self.handle, "h", None):
            raise ValueError("the handle is closed")
        b = ctypes.c_void_p()
        check(_lib.lib().oa_batch_create(self.handle.h, ctypes.byref(b)), "oa_batch_create")
        self.b = b
        self.handle._batches.add(self)
        self._keep = []
        self.count = 0

    def plan(self, psis, outs=None, nkx0=1, nspin=1, tol: float = 1e-15, maxit: int = 1000,
             prec="fp64", S=None, T=None):
        """psis: list of (n_i, 2*npw_i) float64 device tensors; outs / S / T:
        lists of outputs (allocated when None).  nkx0 / nspin / tol / maxit:
        scalars or per-instance lists."""
        count = len(psis)
        per = lambda v, i: v[i] if isinstance(v, (list, tuple)) else v  # noqa: E731
        outs = outs or [torch.empty_like(p) for p in psis]
        S = S or [torch.empty((p.shape[0], p.shape[0]), dtype=torch.float64, device=p.device)
                  for p in psis]
        T = T or [torch.empty((p.shape[0], p.shape[0]), dtype=torch.float64, device=p.device)
                  for p in psis]
        arr = (_lib.oa_instance * max(count, 1))()
        for i, p in enumerate(psis):
            _need(p, f"psis[{i}]")
            _need(outs[i], f"outs[{i}]", p.shape[0])
            arr[i] = _lib.oa_instance(p.shape[0], p.shape[1] // 2, per(nkx0, i), per(nspin, i),
                                      _ptr(p), _ld(p), _ptr(outs[i]), _ld(outs[i]),
                                      _ptr(S[i]), _ptr(T[i]), per(tol, i), per(maxit, i), 0)
        self.handle.use_stream()
        check(_lib.lib().oa_batch_plan(self.b, count, arr, prec_code(prec)), "oa_batch_plan")
        self._keep = [psis, outs, S, T]
        self.psis, self.outs, self.S, self.T = psis, outs, S, T
        self.count = count

    def run(self) -> None:
        self.handle.use_stream()
        check(_lib.lib().oa_batch_run(self.b), "oa_batch_run")

    def launch(self) -> None:
        """Replay the step as a hipGraph (captured on first use after plan)."""
        self.handle.use_stream()
        check(_lib.lib().oa_batch_launch(self.b), "oa_batch_launch")

    def reports(self):
        """[(iterations, done, status, error)] per instance (synchronous)."""
        arr = (_lib.oa_ns_report * max(self.count, 1))()
        check(_lib.lib().oa_batch_reports(self.b, arr, self.count), "oa_batch_reports")
        return [(r.iterations, r.done, r.status, r.error) for r in arr[:self.count]]

    def tune(self, reps: int = 3) -> None:
        self.handle.use_stream()
        check(_lib.lib().oa_batch_tune(self.b, reps), "oa_batch_tune")

    def config(self, i: int):
        t, w = ctypes.c_int(0), ctypes.c_int(0)
        check(_lib.lib().oa_batch_config(self.b, i, ctypes.byref(t), ctypes.byref(w)),
              "oa_batch_config")
        return t.value, w.value

    def table(self) -> str:
        L = _lib.lib()
        need = L.oa_batch_get_table(self.b, None, 0)
        buf = ctypes.create_string_buffer(max(need, 1))
        check(L.oa_batch_get_table(self.b, buf, len(buf)), "oa_batch_get_table")
        return buf.value.decode()

    def set_table(self, text: str) -> None:
        check(_lib.lib().oa_batch_set_table(self.b, text.encode()), "oa_batch_set_table")

    def footprint(self) -> int:
        return int(_lib.lib().oa_batch_footprint(self.b))

    def reserve(self, workspace_bytes: int, table_bytes: int = 1 << 22) -> None:
        """Size the workspace once, so re-plans within it allocate nothing."""
        check(_lib.lib().oa_batch_reserve(self.b, workspace_bytes, table_bytes),
              "oa_batch_reserve")

    def close(self) -> None:
        if getattr(self, "b", None):
            _lib.lib().oa_batch_destroy(self.b)
            self.b = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def instance_footprint(n: int, npw: int, prec="fp64") -> int:
    """Device bytes one instance of (n, npw) holds inside a batch."""
    return int(_lib.lib().oa_instance_footprint(n, npw, prec_code(prec)))


_default: dict[int, Handle] = {}


def default_handle(device: int | None = None) -> Handle:
    d = torch.cuda.current_device() if device is None else device
    h = _default.get(d)
    if h is None:
        h = _default[d] = Handle(d)
    else:
        h.use_stream()
    return h


class PairCalculator:
    """One PairCalculator instance over a whole state block (the reference's
    default decomposition sGrainSize = n collapses the 4-D array to one tile
    per plane; here the plane split is the kernel's internal split-K).

    symmetric=True  -> the psi x psi instance (S, T*psi); alpha = 2 when
                       double packing is on (gamma point), kx=0 weighted.
    symmetric=False -> the psi x F instance (Lambda, alpha = 1).
    complex_math=True -> the CP_PAIRCALC_USES_COMPLEX_MATH build (k-points):
                       complex S = psi psi^H and Lambda = psi F^H
                       (ZGEMM('C','N')), complex T psi / Lambda^H psi; no
                       double packing."""

    def __init__(self, symmetric: bool = True, nkx0: int = 1, double_pack: bool = True,
                 prec="fp64", handle: Handle | None = None, complex_math: bool = False):
        self.symmetric = symmetric
        self.nkx0 = nkx0
        self.double_pack = double_pack and not complex_math
        self.prec = prec
        self.complex_math = complex_math
        self.handle = handle or default_handle()

    def multiplyForward(self, left: torch.Tensor, right: torch.Tensor | None = None,
                        diag: float | None = None):
        """Symmetric: S (and, when `diag` is given, max|S - diag*I| for the
        dynamics tolerance check, returned as (S, maxdev)).  Asymmetric:
        Lambda from (psi, force)."""
        h = self.handle
        if self.complex_math:
            if not self.symmetric:
                if right is None:
                    raise ValueError("asymmetric instance needs the right (force) data")
                return h.zlambda(left, right, prec=self.prec)
            md = None
            if diag is not None:
                md = torch.zeros(1, dtype=torch.float64, device=left.device)
            # Ortho::array_diag_max on the complex S (ortho.h:282-313, myabs = |z|)
            S = h.zoverlap(left, prec=self.prec, diag=diag or 0.0, maxdev=md)
            return S if md is None else (S, float(md.item()))
        if self.symmetric:
            alpha = 2.0 if self.double_pack else 1.0
            md = None
            if diag is not None:
                md = torch.zeros(1, dtype=torch.float64, device=left.device)
            S = h.overlap_sym(left, nkx0=self.nkx0 if self.double_pack else 0, alpha=alpha,
                              prec=self.prec, diag=diag or 0.0, maxdev=md)
            return S if md is None else (S, float(md.item()))
        if right is None:
            raise ValueError("asymmetric instance needs the right (force) data")
        return h.lambda_fwd(left, right, alpha=1.0, prec=self.prec)

    def multiplyResult(self, M: torch.Tensor, psi: torch.Tensor, out=None,
                       alpha: float = 1.0, beta: float = 0.0):
        """Symmetric: out = alpha*T*psi + beta*out (ckPairCalculator.C:1835);
        asymmetric: out = alpha*Lambda^T*psi + beta*out (:1842), Lambda^H
        under complex math."""
        if self.complex_math:
            return self.handle.zbackmul(M, psi, transM=not self.symmetric, alpha=alpha,
                                        beta=beta, out=out, prec=self.prec)
        return self.handle.backmul(M, psi, transM=not self.symmetric, alpha=alpha, beta=beta,
                                   out=out, prec=self.prec)

    def bwMultiplyDynOrthoT(self, orthoT: torch.Tensor, F: torch.Tensor, out=None,
                            beta: float = 0.0):
        """-F*T (ckPairCalculator.C:2018, alpha = -1)."""
        if self.complex_math:
            return self.handle.zbackmul(orthoT, F, alpha=-1.0, beta=beta, out=out,
                                        prec=self.prec)
        return self.handle.backmul(orthoT, F, transM=False, alpha=-1.0, beta=beta, out=out,
                                   prec=self.prec)

    def multiplyPsiV(self, orthoT: torch.Tensor, psiV: torch.Tensor, out=None):
        if self.complex_math:
            return self.handle.zbackmul(orthoT, psiV, out=out, prec=self.prec)
        return self.handle.backmul(orthoT, psiV, transM=False, out=out, prec=self.prec)


class Ortho:
    """The Ortho chare array collapsed onto one GPU (orthoGrainSize = n)."""

    def __init__(self, nspin: int = 1, invsqr_tolerance: float = 1e-15,
                 invsqr_max_iter: int = 1000, max_tolerance: float | None = None,
                 handle: Handle | None = None, complex_math: bool = False,
                 stop_rule: str = "reference"):
        """complex_math: S, T, Lambda are complex128 (k-points).  stop_rule
        (complex only): "reference" = sqrt(Re sum dA^2 / n^2) as the
        reference's complex tolerance_check forms it, "modulus" = sum |dA|^2."""
        if stop_rule not in ("reference", "modulus"):
            raise ValueError(stop_rule)
        self.complex_math = complex_math
        self.stop_rule = stop_rule
        self.nspin = nspin
        self.lsda = nspin != 1
        self.invsqr_tolerance = invsqr_tolerance
        self.invsqr_max_iter = invsqr_max_iter
        self.max_tolerance = max_tolerance
        self.handle = handle or default_handle()
        self.iterations = 0
        self.error = float("nan")

    def start_calc(self, S: torch.Tensor) -> torch.Tensor:
        if self.complex_math:
            T, self.iterations, self.error = self.handle.zs_to_t(
                S, nspin=self.nspin, tol=self.invsqr_tolerance, maxit=self.invsqr_max_iter,
                modulus=self.stop_rule == "modulus")
            return T
        T, self.iterations, self.error = self.handle.s_to_t(
            S, nspin=self.nspin, tol=self.invsqr_tolerance, maxit=self.invsqr_max_iter)
        return T

    @property
    def diag(self) -> float:
        """The diagonal Ortho::array_diag_max measures against (ortho.h:
        282-313): 1 under lsda, else 2.  Pass it to
        PairCalculator.multiplyForward(diag=...) to get max|S - d I| from the
        overlap epilogue."""
        return 1.0 if self.lsda else 2.0

    def tolerance_ok(self, maxdev: float) -> bool:
        """Ortho::maxCheck (ortho.C:455-475): False means recompute PsiV."""
        return self.max_tolerance is None or maxdev < self.max_tolerance

    def acceptSectionLambda(self, L: torch.Tensor) -> torch.Tensor:
        """Minimisation: Lambda (x2 under lsda) goes back to the asymmetric PCs."""
        return L * 2.0 if self.lsda else L

    def gamma(self, orthoT: torch.Tensor, L: torch.Tensor) -> torch.Tensor:
        """Dynamics: Gamma = T * Lambda (CLA multiply alpha=1, beta=0)."""
        n = L.shape[0]
        Lx = L * 2.0 if self.lsda else L
        G = torch.empty_like(L)
        # row-major G = T * Lx  ==  column-major G^T = Lx^T * T^T
        if self.complex_math:
            st = self.handle.zgemm("N", "N", n, n, n, 1.0, Lx, n, orthoT, n, 0.0, G, n)
            check(st, "oa_zgemm")
            return G
        st = self.handle.dgemm("N", "N", n, n, n, 1.0, Lx, n, orthoT, n, 0.0, G, n)
        check(st, "oa_dgemm")
        return G


RAD2I = 1.0 / math.sqrt(2.0)  # as the reference computes it (CP_State_GSpacePlane.C:1929)


class GSpacePlane:
    """GSpace side of the Lambda round trip for a whole state block
    (CP_State_GSpacePlane.C).  The accumulation of the PairCalculator's
    backward result into the force runs fused into the backward multiply
    (oa_lambda_force), so the result block the reference ships back
    (partialResultMsg / CkReductionMsg) is never materialised.

    cp_min_opt = 1: minimisation, 0: dynamics.  double_pack: gamma-point
    storage, where the leading nkx0 points are the kx = 0 plane."""

    def __init__(self, nkx0: int = 1, double_pack: bool = True, cp_min_opt: int = 1,
                 nspin: int = 1, prec="fp64", handle: Handle | None = None):
        self.nkx0 = nkx0 if double_pack else 0
        self.double_pack = double_pack
        self.cp_min_opt = cp_min_opt
        self.lsda = nspin != 1
        self.prec = prec
        self.handle = handle or default_handle()

    def _kx0_points(self, X: torch.Tensor) -> int:
        # single pack scales every point (CP_State_GSpacePlane.C:1942-1946)
        return self.nkx0 if self.double_pack else X.shape[1] // 2

    def scaleForLambda(self, psi: torch.Tensor, force: torch.Tensor) -> None:
        """Dynamics, before the Lambda forward (CP_State_GSpacePlane.C:
        1926-1947): psi *= 1/sqrt2 and force *= sqrt2 on the kx = 0 points
        (double pack) or on every point (single pack), in place."""
        if self.cp_min_opt != 0:
            return
        k = self._kx0_points(psi)
        self.handle.scale_kx0(psi, k, RAD2I)
        self.handle.scale_kx0(force, k, math.sqrt(2.0))

    def unpackLambda(self, L: torch.Tensor, psi: torch.Tensor, force: torch.Tensor,
                     reduction: bool = False) -> torch.Tensor:
        """Minimisation: force -= w .* (Psi Lambda), the asymmetric backward
        product (ckPairCalculator.C:1842, out_j = sum_i psi_i L[i][j])
        accumulated with the GSpace weights.
          partialResultMsg (CP_State_GSpacePlane.C:2168-2195): w = 1/2 at
            kx = 0 and 1 elsewhere (double pack); 1/2 everywhere (single pack).
          CkReductionMsg (:2068-2106): w = 1/ov at kx = 0, 2/ov elsewhere
            (double pack), 1/ov everywhere (single pack); ov = 2 (lda) or 1
            (lsda).
        In place; returns force."""
        if self.cp_min_opt != 1:
            raise ValueError("unpackLambda accumulates minimisation forces; dynamics uses "
                             "unpackGamma")
        if reduction:
            ov = 1.0 if self.lsda else 2.0
            alpha, wk = (-2.0 / ov, 0.5) if self.double_pack else (-1.0 / ov, 1.0)
        else:
            alpha, wk = (-1.0, 0.5) if self.double_pack else (-0.5, 1.0)
        return self.handle.lambda_force(L, psi, force, transM=True, nkx0=self.nkx0, wkx0=wk,
                                        alpha=alpha, beta=1.0, prec=self.prec)

    def unpackGamma(self, gamma: torch.Tensor, orthoT: torch.Tensor, psi: torch.Tensor,
                    F: torch.Tensor, force: torch.Tensor) -> torch.Tensor:
        """Dynamics: the PairCalculator returns data = Psi Gamma^T - F T
        (multiplyResult with Gamma accumulated onto bwMultiplyDynOrthoT's
        -F T, ckPairCalculator.C:1965-2031, :1312-1316); the GSpace sets
        force = -data (CP_State_GSpacePlane.C:2177-2183) and doLambda
        scales its kx = 0 points by 1/sqrt2 (:2240-2245).  psi and F are the
        blocks the PairCalculator received (after scaleForLambda); force
        must not share memory with F.  Returns force."""
        if self.cp_min_opt != 0:
            raise ValueError("unpackGamma is the dynamics path")
        if force.untyped_storage().data_ptr() == F.untyped_storage().data_ptr():
            raise ValueError("force must not alias F: F T is read while force is written")
        wk = RAD2I if self.nkx0 > 0 else 1.0
        h = self.handle
        h.lambda_force(gamma, psi, force, transM=True, nkx0=self.nkx0, wkx0=wk, alpha=-1.0,
                       beta=0.0, prec=self.prec)
        # -F T as bwMultiplyDynOrthoT forms it: DGEMM('N','N') with T as received
        return h.lambda_force(orthoT, F, force, transM=False, nkx0=self.nkx0, wkx0=wk,
                              alpha=1.0, beta=1.0, prec=self.prec)
