"""ZGEMM drop-in (complex-math PairCalculator GEMM) against the oracle's
restatement of zgemm.f, which is pinned bitwise to the compiled reference.
Tolerance: 1e-13 of max|C| (the GPU forms four real FP64-MFMA products and
combines them; zgemm.f accumulates complex products in order)."""
import ctypes
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd import _lib  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
ZCASES = [(1 + 0j, 0j), (0.5 - 1.5j, 0.25 + 2j), (0j, 1.5j), (0j, 0j), (2j, 1 + 0j)]


def cn(rng, *shape):
    return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)


def zgemm_host(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C, ldc):
    L = _lib.lib()
    f = L.zgemm_
    f.restype = None
    ci = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    al = np.array([alpha], dtype=np.complex128)
    be = np.array([beta], dtype=np.complex128)
    f(ctypes.c_char_p(ta.encode()), ctypes.c_char_p(tb.encode()), ctypes.byref(ci(m)),
      ctypes.byref(ci(n)), ctypes.byref(ci(k)), al.ctypes.data_as(dp), A.ctypes.data_as(dp),
      ctypes.byref(ci(lda)), B.ctypes.data_as(dp), ctypes.byref(ci(ldb)), be.ctypes.data_as(dp),
      C.ctypes.data_as(dp), ctypes.byref(ci(ldc)))


def close(got, want):
    scale = max(np.abs(want).max(), 1.0)
    return np.abs(got - want).max() <= 1e-13 * scale


@pytest.mark.parametrize("ta", list("NTC"))
@pytest.mark.parametrize("tb", list("NTC"))
def test_zgemm_host_pointers(oracle, ta, tb):
    rng = np.random.default_rng(ord(ta) * 5 + ord(tb))
    m, n, k = 37, 29, 301
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ra + 2, rb + 1, m + 3
    A, B, C0 = cn(rng, ca, lda), cn(rng, cb, ldb), cn(rng, n, ldc)
    for alpha, beta in ZCASES:
        want, got = C0.copy(), C0.copy()
        oracle.zgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, want, ldc)
        zgemm_host(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, got, ldc)
        assert close(got, want), (ta, tb, alpha, beta, np.abs(got - want).max())
        # rows between m and ldc are not part of C
        assert np.array_equal(got[:, m:], C0[:, m:])


def test_zgemm_device_pointers(handle, oracle):
    rng = np.random.default_rng(9)
    m, n, k = 64, 48, 1000
    # op(A) = A^H with A stored k x m (lda = k), B stored k x n, C m x n
    A, B, C0 = cn(rng, m, k), cn(rng, n, k), cn(rng, n, m)
    want = C0.copy()
    oracle.zgemm_cm("C", "N", m, n, k, 0.5 - 1j, A, k, B, k, 2.0, want, m)
    dA, dB = torch.from_numpy(A).cuda(), torch.from_numpy(B).cuda()
    dC = torch.from_numpy(C0.copy()).cuda()
    info = handle.zgemm("C", "N", m, n, k, 0.5 - 1j, dA, k, dB, k, 2.0, dC, m)
    torch.cuda.synchronize()
    assert info == 0
    assert close(dC.cpu().numpy(), want)
    assert handle.zgemm("C", "N", m, n, k, 1, dA, k - 1, dB, k, 0, dC, m) == 8


@pytest.mark.parametrize("kpt", [0, 1])
def test_zgemm_gwbse_states(kpt):
    """The reference's own complex states: S = Psi^H Psi through the drop-in
    matches the reference ZGEMM's S and is the identity to 1e-14."""
    d = np.load(os.path.join(GOLDEN, f"gwbse_kpt{kpt}.npz"))
    psi = np.ascontiguousarray(d["psi"])
    nst, ncoef = psi.shape
    S = np.zeros((nst, nst), dtype=np.complex128)
    zgemm_host("C", "N", nst, nst, ncoef, 1.0, psi, ncoef, psi, ncoef, 0.0, S, nst)
    assert np.abs(S - d["S_cm"]).max() < 1e-14
    assert np.abs(S.T - np.eye(nst)).max() < 1e-14
