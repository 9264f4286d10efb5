"""GPU parity of the FP64-MFMA path against the CPU oracle (through the C ABI).

Tolerances: FP64 products accumulate in a different order than the
reference's sequential DGEMM dot products, so results agree to rounding, not
bitwise.  S, T and T*psi are held to 1e-12 relative to the largest entry
(the bound the plain FP64 form promises); end-to-end orthonormality to 1e-12.
"""
import ctypes
import glob
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd import _lib  # noqa: E402
from openatom_amd.workload import make_force, make_states  # noqa: E402

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
DEV = "cuda"


def rel(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def cm(x):  # numpy column-major flat buffer -> device tensor
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("m,n,k", [(16, 16, 4), (70, 33, 129), (128, 256, 64), (5, 300, 17)])
def test_dgemm_device(handle, oracle, ta, tb, m, n, k):
    rng = np.random.default_rng(m * 7 + n + k)
    lda = (m if ta == "N" else k) + 1
    ldb = (k if tb == "N" else n) + 3
    ldc = m + 2
    A = rng.standard_normal(lda * (k if ta == "N" else m))
    B = rng.standard_normal(ldb * (n if tb == "N" else k))
    C0 = rng.standard_normal(ldc * n)
    for alpha, beta in [(1.0, 0.0), (-0.5, 2.0)]:
        want = C0.copy()
        oracle.dgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, want, ldc)
        dA, dB, dC = cm(A), cm(B), cm(C0)
        info = handle.dgemm(ta, tb, m, n, k, alpha, dA, lda, dB, ldb, beta, dC, ldc)
        assert info == 0
        torch.cuda.synchronize()
        got = dC.cpu().numpy()
        assert rel(got, want) < 1e-13, (ta, tb, m, n, k, alpha, beta)
        # padding rows between m and ldc untouched
        assert np.array_equal(got.reshape(n, ldc)[:, m:], C0.reshape(n, ldc)[:, m:])


def test_dgemm_drop_in_host_pointers(oracle):
    """The Fortran-ABI dgemm_ of liboa_ortho.so on host buffers, as a Charm++
    PairCalculator would call it (ckPairCalculator.C:889)."""
    L = _lib.lib()
    f = L.dgemm_
    f.restype = None
    rng = np.random.default_rng(3)
    nX, nY, K = 40, 24, 2 * 515
    left = rng.standard_normal((nX, K))
    right = rng.standard_normal((nY, K))
    out = np.zeros(nX * nY)
    ci, cd = ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    args = [ctypes.c_char_p(b"T"), ctypes.c_char_p(b"N"), ctypes.byref(ci(nY)),
            ctypes.byref(ci(nX)), ctypes.byref(ci(K)), ctypes.byref(cd(2.0)),
            right.ctypes.data_as(dp), ctypes.byref(ci(K)), left.ctypes.data_as(dp),
            ctypes.byref(ci(K)), ctypes.byref(cd(0.0)), out.ctypes.data_as(dp),
            ctypes.byref(ci(nY))]
    f(*args)
    want = np.zeros(nX * nY)
    oracle.dgemm_cm("T", "N", nY, nX, K, 2.0, right.ravel(), K, left.ravel(), K, 0.0, want, nY)
    assert rel(out, want) < 1e-13
    # XERBLA semantics: invalid LDC -> silent return, C untouched
    out2 = np.full(nX * nY, 3.0)
    args[11] = out2.ctypes.data_as(dp)
    args[12] = ctypes.byref(ci(nY - 1))
    f(*args)
    assert np.all(out2 == 3.0)


@pytest.mark.parametrize("n,npw,nkx0", [(128, 8192, 1), (37, 1000, 3), (256, 2048, 1),
                                        (1, 17, 1), (130, 700, 0), (200, 4096, 64)])
def test_overlap_vs_oracle(handle, oracle, n, npw, nkx0):
    psi = make_states(n, npw, eps=1e-3, seed=n + npw, nkx0=max(nkx0, 1)).numpy()
    want = oracle.overlap(psi, nkx0=nkx0, nplanes=4)
    S = handle.overlap_sym(torch.from_numpy(psi).to(DEV), nkx0=nkx0, alpha=2.0)
    got = S.cpu().numpy()
    assert rel(got, want) < 1e-12
    assert np.array_equal(got, got.T)


def test_overlap_strided_rows_and_maxdev(handle, oracle):
    n, npw = 96, 1500
    psi = make_states(n, npw, eps=1e-2, seed=21).numpy()
    big = np.zeros((n, 2 * npw + 10))
    big[:, : 2 * npw] = psi
    dbig = torch.from_numpy(big).to(DEV)
    md = torch.zeros(1, dtype=torch.float64, device=DEV)
    S = handle.overlap_sym(dbig[:, : 2 * npw], nkx0=1, alpha=2.0, diag=2.0, maxdev=md)
    want = oracle.overlap(psi, 1)
    assert rel(S.cpu().numpy(), want) < 1e-12
    dev = np.abs(want - 2.0 * np.eye(n)).max()
    assert abs(md.item() - dev) <= 1e-12 * max(dev, 1.0)


def test_lambda_vs_oracle(handle, oracle):
    n, npw = 150, 3000
    psi = make_states(n, npw, eps=1e-3, seed=31).numpy()
    F = make_force(n, npw, seed=32).numpy()
    want = oracle.lambda_fwd(psi, F)
    L = handle.lambda_fwd(torch.from_numpy(psi).to(DEV), torch.from_numpy(F).to(DEV))
    assert rel(L.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("n,nspin", [(128, 2), (128, 1), (300, 2), (7, 1)])
def test_s_to_t_vs_oracle(handle, oracle, n, nspin):
    psi = make_states(n, max(4 * n, 64), eps=1e-2, seed=40 + n).numpy()
    S = oracle.overlap(psi, 1)
    if nspin == 1:
        S = 2.0 * S
    Tw, itw, errw, stw = oracle.s_to_t(S, nspin=nspin, tol=1e-13, maxit=100)
    T, it, err = handle.s_to_t(torch.from_numpy(S).to(DEV), nspin=nspin, tol=1e-13, maxit=100)
    assert rel(T.cpu().numpy(), Tw) < 1e-12
    assert abs(it - itw) <= 1 and err <= 1e-13


def test_s_to_t_iteration_limit_raises(handle, oracle):
    psi = make_states(64, 512, eps=1e-2, seed=50).numpy()
    S = torch.from_numpy(oracle.overlap(psi, 1)).to(DEV)
    with pytest.raises(_lib.OrthoNotConverged):
        handle.s_to_t(S, nspin=2, tol=1e-15, maxit=2)
    assert handle.last_iterations == 2
    T, it, err = handle.s_to_t(S, nspin=2, tol=1e-15, maxit=100)
    with pytest.raises(_lib.OrthoNotConverged):  # converged exactly at maxit still aborts
        handle.s_to_t(S, nspin=2, tol=1e-15, maxit=it)


def test_s_to_t_async_matches_sync(handle, oracle):
    psi = make_states(100, 800, eps=1e-2, seed=51).numpy()
    S = torch.from_numpy(oracle.overlap(psi, 1)).to(DEV)
    T1, it1, err1 = handle.s_to_t(S, nspin=2, tol=1e-14, maxit=50)
    T2 = torch.empty_like(S)
    rep = torch.zeros(3, dtype=torch.float64, device=DEV)  # 24 bytes
    handle.s_to_t_async(S, T2, rep, nspin=2, tol=1e-14, maxit=12)
    torch.cuda.synchronize()
    r = _lib.oa_ns_report.from_buffer_copy(rep.cpu().numpy().tobytes())
    assert r.done == 1 and r.status == 0 and r.iterations == it1
    assert torch.equal(T1, T2)


@pytest.mark.parametrize("transM", [False, True])
def test_backmul_vs_oracle(handle, oracle, transM):
    rng = np.random.default_rng(60)
    n, npw = 160, 2500
    M = rng.standard_normal((n, n))
    psi = rng.standard_normal((n, 2 * npw))
    out0 = rng.standard_normal((n, 2 * npw))
    want = oracle.backmul(M, psi, transT=transM, alpha=-1.0, beta=0.5, out=out0.copy())
    got = handle.backmul(torch.from_numpy(M).to(DEV), torch.from_numpy(psi).to(DEV),
                         transM=transM, alpha=-1.0, beta=0.5,
                         out=torch.from_numpy(out0).to(DEV))
    assert rel(got.cpu().numpy(), want) < 1e-12


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "step_*.npz"))))
def test_golden_fixtures_on_gpu(handle, path):
    g = np.load(path)
    psi = torch.from_numpy(g["psi"]).to(DEV)
    nkx0 = int(g["nkx0"])
    for nspin in (1, 2):
        S, T, out, it, err = handle.ortho_step(psi, nkx0=nkx0, nspin=nspin, tol=1e-15,
                                               maxit=1000)
        assert rel(S.cpu().numpy(), g[f"S_nspin{nspin}"]) < 1e-12
        assert rel(T.cpu().numpy(), g[f"T_nspin{nspin}"]) < 1e-12
        assert rel(out.cpu().numpy(), g[f"psi_out_nspin{nspin}"]) < 1e-12
        assert abs(it - int(g[f"iters_nspin{nspin}"])) <= 1
    L = handle.lambda_fwd(psi, torch.from_numpy(g["F"]).to(DEV))
    assert rel(L.cpu().numpy(), g["L"]) < 1e-12
    LT = handle.backmul(L, psi, transM=True)
    assert rel(LT.cpu().numpy(), g["LT_psi"]) < 1e-12


def test_config_a_step_vs_oracle(handle, oracle):
    """Config A end to end: n=128, npw=8192, eps=1e-3, seed 0, tol 1e-12."""
    psi = make_states(128, 8192, eps=1e-3, seed=0).numpy()
    r = oracle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=1000, nplanes=8)
    S, T, out, it, err = handle.ortho_step(torch.from_numpy(psi).to(DEV), nkx0=1, nspin=2,
                                           tol=1e-12, maxit=1000)
    assert rel(S.cpu().numpy(), r["S"]) < 1e-12
    assert rel(T.cpu().numpy(), r["T"]) < 1e-12
    assert rel(out.cpu().numpy(), r["psi_out"]) < 1e-12
    res = oracle.orthonormality_residual(out.cpu().numpy(), 1, occ=1.0, nplanes=8)
    assert res < 1e-12


def test_config_b_properties_full_size(handle):
    """Config B size (n=512, npw=65536) through size-independent properties:
    S exactly symmetric, T (S) T = I and the orthonormality of T psi,
    checked with torch's own FP64 GEMMs (an independent implementation)."""
    n, npw = 512, 65536
    psi = make_states(n, npw, eps=1e-3, seed=0, device=DEV)
    S, T, out, it, err = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=100)
    assert torch.equal(S, S.T)
    w = torch.full((2 * npw,), 2.0, dtype=torch.float64, device=DEV)
    w[:2] = 1.0
    S_t = (psi * w) @ psi.T
    assert (S - S_t).abs().max().item() < 1e-12 * S_t.abs().max().item()
    I = torch.eye(n, dtype=torch.float64, device=DEV)
    assert (T @ S @ T - I).abs().max().item() < 1e-12
    S2 = (out * w) @ out.T
    assert (S2 - I).abs().max().item() < 1e-12
