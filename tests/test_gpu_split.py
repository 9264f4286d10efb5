"""Split-precision (Ozaki digit slice) path against the FP64 path.

The FP64 path is itself pinned to the oracle at 1e-12 (test_gpu_f64.py), so
it serves as the reference here.  Tolerances follow the digit arithmetic:
a leading digit of |a_1| <= 127 and S-1 digits in base 254 leave a
representation error of 2^-8 254^-(S-1) of the chunk maximum per element
(measured |dS| / max|S|: x3 1.4e-6, x4 2.6e-8, x5 2.9e-11, x6 4.4e-13); the
bounds below keep a margin of about 10.  Digit sums are exact integers, so
int8, fp16 and bf16 digits must agree to FP64 rounding of the final
combination.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd import _lib  # noqa: E402
from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"
TOL = {2: 1e-3, 3: 1e-5, 4: 3e-7, 5: 3e-10, 6: 5e-12}


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.fixture(scope="module")
def block():
    psi = make_states(256, 16384, eps=1e-3, seed=3, device=DEV, rng="torch")
    return psi


@pytest.mark.parametrize("prec", ["int8x3", "int8x4", "int8x5", "int8x6", "fp16x3", "bf16x3",
                                  "fp16x5", "bf16x5", "fp32x5"])
def test_split_overlap_accuracy(handle, block, prec):
    S64 = handle.overlap_sym(block, nkx0=1, alpha=2.0, prec="fp64")
    Ss = handle.overlap_sym(block, nkx0=1, alpha=2.0, prec=prec)
    torch.cuda.synchronize()
    e = rel(Ss, S64)
    slices = int(prec.split("x")[1])
    print(f"{prec}: overlap max rel error {e:.3e}")
    assert e < TOL[slices], e
    assert torch.equal(Ss, Ss.T)


@pytest.mark.parametrize("slices", [3, 4])
def test_split_digit_types_agree(handle, block, slices):
    a = handle.overlap_sym(block, nkx0=1, prec=f"int8x{slices}")
    b = handle.overlap_sym(block, nkx0=1, prec=f"fp16x{slices}")
    c = handle.overlap_sym(block, nkx0=1, prec=f"bf16x{slices}")
    torch.cuda.synchronize()
    assert rel(b, a) < 1e-14 and rel(c, a) < 1e-14


@pytest.mark.parametrize("slices", [3, 5])
def test_split_fp32_digits_agree(handle, block, slices):
    """fp32 digits on v_mfma_f32_32x32x2_f32: the same exact level sums."""
    a = handle.overlap_sym(block, nkx0=1, prec=f"int8x{slices}")
    d = handle.overlap_sym(block, nkx0=1, prec=f"fp32x{slices}")
    torch.cuda.synchronize()
    assert rel(d, a) < 1e-14


@pytest.mark.parametrize("prec", ["int8x4", "int8x5", "int8x6l", "fp16x5", "bf16x4", "fp32x5"])
def test_split_backmul_accuracy(handle, block, prec):
    n = block.shape[0]
    S = handle.overlap_sym(block, nkx0=1, prec="fp64")
    T, it, err = handle.s_to_t(S, nspin=2, tol=1e-13, maxit=50)
    o64 = handle.backmul(T, block, prec="fp64")
    os_ = handle.backmul(T, block, prec=prec)
    torch.cuda.synchronize()
    e = rel(os_, o64)
    print(f"{prec}: back-multiply max rel error {e:.3e}")
    slices = int(prec.split("x")[1].rstrip("l"))
    assert e < 10 * TOL[slices - (1 if prec.endswith("l") else 0)], e
    # transposed operand and alpha/beta
    M = torch.randn(n, n, dtype=torch.float64, device=DEV)
    base = torch.randn_like(block)
    w64 = handle.backmul(M, block, transM=True, alpha=-0.5, beta=2.0, out=base.clone(),
                         prec="fp64")
    ws = handle.backmul(M, block, transM=True, alpha=-0.5, beta=2.0, out=base.clone(), prec=prec)
    ew = rel(ws, w64)
    print(f"{prec}: general M^T back-multiply max rel error {ew:.3e}")
    assert ew < 1e4 * TOL[slices]


def test_split_reuse_digits_identical(handle, block):
    S = handle.overlap_sym(block, nkx0=1, prec="int8x5")
    T, it, err = handle.s_to_t(S, nspin=2, tol=1e-13, maxit=50)
    a = handle.backmul(T, block, prec="int8x5")
    handle.overlap_sym(block, nkx0=1, prec="int8x5")
    b = handle.backmul(T, block, prec="int8x5", reuse_digits=True)
    torch.cuda.synchronize()
    assert torch.equal(a, b)


@pytest.mark.parametrize("n,npw,nkx0", [(37, 1000, 3), (130, 700, 0), (1, 64, 1), (200, 5000, 2)])
def test_split_ragged_sizes(handle, n, npw, nkx0):
    psi = make_states(n, npw, eps=1e-2, seed=n, nkx0=max(nkx0, 1), device=DEV, rng="torch")
    S64 = handle.overlap_sym(psi, nkx0=nkx0, prec="fp64")
    Ss = handle.overlap_sym(psi, nkx0=nkx0, prec="int8x5")
    assert rel(Ss, S64) < TOL[5]
    # a near-identity T (the path's case) and a general matrix (no diagonal gain)
    X = 1e-3 * torch.randn(n, n, dtype=torch.float64, device=DEV)
    T = torch.eye(n, dtype=torch.float64, device=DEV) + (X + X.T)
    assert rel(handle.backmul(T, psi, prec="int8x5"), handle.backmul(T, psi, prec="fp64")) < 1e-9
    M = torch.randn(n, n, dtype=torch.float64, device=DEV)
    assert rel(handle.backmul(M, psi, prec="int8x5"), handle.backmul(M, psi, prec="fp64")) < 1e-6


@pytest.mark.parametrize("prec", ["int8x5", "fp16x5", "fp32x5", "int8x6l"])
def test_split_step_meets_residual_target(handle, prec):
    """n=512, npw=65536 (the plain-vs-split configuration): one split step
    leaves the orthonormality residual <= 1e-10 and T within 1e-10 of FP64."""
    n, npw = 512, 65536
    psi = make_states(n, npw, eps=1e-3, seed=0, device=DEV, rng="torch")
    S64, T64, o64, it64, _ = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=50)
    Ss, Ts, os_, its, _ = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=50,
                                            prec=prec)
    eye = torch.eye(n, dtype=torch.float64, device=DEV)
    res = (handle.overlap_sym(os_, nkx0=1, prec="fp64") - eye).abs().max().item()
    res64 = (handle.overlap_sym(o64, nkx0=1, prec="fp64") - eye).abs().max().item()
    print(f"residual {prec} {res:.3e}  fp64 {res64:.3e}  |dT| {(Ts - T64).abs().max().item():.3e}")
    assert res <= 1e-10
    assert (Ts - T64).abs().max().item() <= 1e-10


@pytest.mark.parametrize("prec", ["int8x5", "fp16x5", "int8x3"])
def test_split_lambda(handle, prec):
    """Asymmetric Lambda = psi F^T on digits of psi and F (own chunk
    exponents), all tiles, against the FP64 product; F has a different scale
    from psi (std 0.1) and no orthogonality."""
    from openatom_amd.workload import make_force
    psi = make_states(200, 6000, eps=1e-3, seed=5, device=DEV, rng="torch")
    F = make_force(200, 6000, seed=6, device=DEV, rng="torch")
    L64 = handle.lambda_fwd(psi, F, alpha=1.0, prec="fp64")
    Ls = handle.lambda_fwd(psi, F, alpha=1.0, prec=prec)
    torch.cuda.synchronize()
    e = rel(Ls, L64)
    print(f"{prec}: Lambda max rel error {e:.3e}")
    slices = int(prec.split("x")[1])
    assert e < 10 * TOL[slices]
