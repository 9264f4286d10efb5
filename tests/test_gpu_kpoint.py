"""GPU parity of the complex-math (k-point) path -- oa_zoverlap, oa_zlambda,
oa_zs_to_t, oa_zbackmul, oa_zortho_step and the complex PairCalculator /
Ortho mirrors -- against the CPU oracle (itself pinned by numpy's complex
products and Hermitian eigensolver, the real oracle and the reference's
k-point state files).

Tolerances: the FP64-MFMA products accumulate in another order than the
reference's ZGEMM loops, so results agree to 1e-12 relative to the largest
entry; Hermitian symmetry of S is exact (upper triangle mirrored).
"""
import os

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.ortho import Ortho, PairCalculator  # noqa: E402
from openatom_amd.workload import make_kpoint_states, make_states  # noqa: E402

DEV = "cuda"
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def cplx(psi):
    return psi[:, 0::2] + 1j * psi[:, 1::2]


@pytest.mark.parametrize("n,npw", [(64, 2048), (37, 1000), (200, 5000), (1, 64)])
def test_zoverlap_and_zlambda(handle, oracle, n, npw):
    psi_h = make_kpoint_states(n, npw, eps=1e-2, seed=n).numpy()
    F_h = make_kpoint_states(n, npw, eps=1e-1, seed=n + 1).numpy()
    S = handle.zoverlap(dev(psi_h))
    L = handle.zlambda(dev(psi_h), dev(F_h), alpha=-0.5)
    torch.cuda.synchronize()
    S_h, L_h = S.cpu().numpy(), L.cpu().numpy()
    assert rel(S_h, oracle.zoverlap(psi_h)) < 1e-13
    assert np.array_equal(S_h, S_h.conj().T)
    assert rel(L_h, -0.5 * oracle.zlambda(psi_h, F_h)) < 1e-13


@pytest.mark.parametrize("kpt", [0, 1])
def test_zoverlap_gwbse_fixture(handle, kpt):
    d = np.load(os.path.join(GOLDEN, f"gwbse_kpt{kpt}.npz"))
    psi = np.ascontiguousarray(d["psi"]).view(np.float64)
    S = handle.zoverlap(dev(psi)).cpu().numpy()
    assert np.abs(S - np.eye(S.shape[0])).max() < 5e-15


@pytest.mark.parametrize("nspin,modulus", [(1, False), (1, True), (2, True)])
def test_zs_to_t_vs_oracle(handle, oracle, nspin, modulus):
    n = 96
    psi_h = make_kpoint_states(n, 3000, eps=2e-2, seed=5).numpy()
    S_h = oracle.zoverlap(psi_h)
    if nspin == 1:
        S_h = 2.0 * S_h
    Tw, itw, errw, stw = oracle.zs_to_t(S_h, nspin=nspin, tol=1e-14, maxit=60, modulus=modulus)
    T, it, err = handle.zs_to_t(dev(S_h), nspin=nspin, tol=1e-14, maxit=60, modulus=modulus)
    assert stw == 0 and it == itw
    assert rel(T.cpu().numpy(), Tw) < 1e-12
    if modulus:  # the reference rule may stop early on a NaN error
        B = S_h / 2.0 if nspin == 1 else S_h
        Th = T.cpu().numpy()
        assert np.abs(Th @ B @ Th - np.eye(n)).max() < 1e-12


def test_zs_to_t_reference_rule_under_lsda(handle, oracle):
    """The reference's complex stop rule ends after one iteration with a NaN
    error on this block, on the device as in the oracle."""
    psi_h = make_kpoint_states(30, 400, eps=1e-2, seed=8).numpy()
    S_h = oracle.zoverlap(psi_h)
    Tw, itw, errw, _ = oracle.zs_to_t(S_h, nspin=2, tol=1e-15, maxit=50)
    T, it, err = handle.zs_to_t(dev(S_h), nspin=2, tol=1e-15, maxit=50)
    assert itw == it == 1 and np.isnan(err) and np.isnan(errw)
    assert rel(T.cpu().numpy(), Tw) < 1e-13


def test_zs_to_t_real_input_matches_real_path(handle):
    psi = make_states(64, 2000, eps=1e-2, seed=9, device=DEV, rng="torch")
    S = handle.overlap_sym(psi, nkx0=1, prec="fp64")
    Tr, itr, _ = handle.s_to_t(S, nspin=2, tol=1e-14, maxit=50)
    Tz, itz, _ = handle.zs_to_t(S.to(torch.complex128), nspin=2, tol=1e-14, maxit=50)
    torch.cuda.synchronize()
    assert itr == itz
    assert rel(Tz.real.cpu().numpy(), Tr.cpu().numpy()) < 1e-13
    assert not Tz.imag.any().item()


@pytest.mark.parametrize("transM", [False, True])
@pytest.mark.parametrize("n,npw", [(80, 3000), (33, 513)])
def test_zbackmul_vs_oracle(handle, oracle, transM, n, npw):
    rng = np.random.default_rng(n)
    psi_h = make_kpoint_states(n, npw, eps=0.0, seed=n).numpy()
    M_h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    out0 = rng.standard_normal(psi_h.shape)
    want = oracle.zbackmul(M_h, psi_h, transM=transM, alpha=-0.5, beta=2.0, out=out0.copy())
    got = handle.zbackmul(dev(M_h), dev(psi_h), transM=transM, alpha=-0.5, beta=2.0,
                          out=dev(out0))
    torch.cuda.synchronize()
    assert rel(got.cpu().numpy(), want) < 1e-13


@pytest.mark.parametrize("nspin", [1, 2])
def test_zortho_step_vs_oracle(handle, oracle, nspin):
    n, npw = 256, 4096
    psi_h = make_kpoint_states(n, npw, eps=1e-2, seed=10).numpy()
    want = oracle.zortho_step(psi_h, nspin=nspin, tol=1e-14, maxit=50, modulus=True)
    S, T, out, it, err = handle.zortho_step(dev(psi_h), nspin=nspin, tol=1e-14, maxit=50,
                                            modulus=True)
    torch.cuda.synchronize()
    assert it == want["iters"]
    assert rel(S.cpu().numpy(), want["S"]) < 1e-13
    assert rel(T.cpu().numpy(), want["T"]) < 1e-12
    assert rel(out.cpu().numpy(), want["psi_out"]) < 1e-12
    c = cplx(out.cpu().numpy())
    occ = 2.0 if nspin == 1 else 1.0
    assert np.abs(c @ c.conj().T - occ * np.eye(n)).max() < 1e-12


def test_complex_pair_calculator_round_trip(handle, oracle):
    """The complex mirrors chained as the reference's instance set runs them:
    S, Ortho's T, T psi; Lambda, Gamma = T Lambda, Lambda^H psi."""
    n, npw = 48, 1500
    psi_h = make_kpoint_states(n, npw, eps=5e-3, seed=11).numpy()
    F_h = make_kpoint_states(n, npw, eps=1e-1, seed=12).numpy()
    psi, F = dev(psi_h), dev(F_h)
    sym = PairCalculator(symmetric=True, complex_math=True, handle=handle)
    asym = PairCalculator(symmetric=False, complex_math=True, handle=handle)
    ortho = Ortho(nspin=1, complex_math=True, stop_rule="modulus", invsqr_tolerance=1e-14,
                  invsqr_max_iter=50, handle=handle)
    S, md = sym.multiplyForward(psi, diag=1.0)
    T = ortho.start_calc(2.0 * S)
    out = sym.multiplyResult(T, psi)
    L = asym.multiplyForward(psi, F)
    G = ortho.gamma(T, L)
    back = asym.multiplyResult(L, psi)
    torch.cuda.synchronize()
    S_h = oracle.zoverlap(psi_h)
    assert abs(md - np.abs(S_h - np.eye(n)).max()) < 1e-14
    Tw, _, _, _ = oracle.zs_to_t(2.0 * S_h, nspin=1, tol=1e-14, maxit=50, modulus=True)
    assert rel(T.cpu().numpy(), Tw) < 1e-12
    assert rel(out.cpu().numpy(), oracle.zbackmul(Tw, psi_h)) < 1e-12
    L_h = oracle.zlambda(psi_h, F_h)
    assert rel(L.cpu().numpy(), L_h) < 1e-13
    assert rel(G.cpu().numpy(), Tw @ L_h) < 1e-12
    assert rel(back.cpu().numpy(), oracle.zbackmul(L_h, psi_h, transM=True)) < 1e-12


# ---- split precision --------------------------------------------------------

@pytest.mark.parametrize("prec", ["int8x5", "fp16x5"])
def test_zortho_step_split_meets_residual(handle, prec):
    """The complex step on digit products: S within 1e-10 of the FP64 step,
    T psi orthonormal to 1e-10 (S->T is FP64 in every mode)."""
    n, npw = 256, 8192
    psi = dev(make_kpoint_states(n, npw, eps=1e-3, seed=13).numpy())
    S64, T64, o64, it64, _ = handle.zortho_step(psi, nspin=2, tol=1e-13, maxit=50, modulus=True)
    Ss, Ts, os_, its, _ = handle.zortho_step(psi, nspin=2, tol=1e-13, maxit=50, modulus=True,
                                             prec=prec)
    torch.cuda.synchronize()
    eS = rel(Ss.cpu().numpy(), S64.cpu().numpy())
    eO = rel(os_.cpu().numpy(), o64.cpu().numpy())
    S2 = handle.zoverlap(os_)
    resid = (S2 - torch.eye(n, dtype=S2.dtype, device=DEV)).abs().max().item()
    print(f"{prec}: S err {eS:.2e}, T psi err {eO:.2e}, residual {resid:.2e}")
    assert np.array_equal(Ss.cpu().numpy(), Ss.cpu().numpy().conj().T)
    assert eS < 1e-10 and eO < 1e-10 and resid < 1e-10


@pytest.mark.parametrize("transM", [False, True])
def test_zbackmul_and_zlambda_split(handle, transM):
    n, npw = 128, 4096
    psi = dev(make_kpoint_states(n, npw, eps=0.0, seed=14).numpy())
    F = dev(make_kpoint_states(n, npw, eps=1e-1, seed=15).numpy())
    rng = np.random.default_rng(16)
    M = dev(rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n)))
    a = handle.zbackmul(M, psi, transM=transM, alpha=-0.5, prec="fp64")
    b = handle.zbackmul(M, psi, transM=transM, alpha=-0.5, prec="int8x5")
    La = handle.zlambda(psi, F, prec="fp64")
    Lb = handle.zlambda(psi, F, prec="int8x5")
    torch.cuda.synchronize()
    # a general matrix gets no diagonal gain: the 4-slice truncation of its digits
    assert rel(b.cpu().numpy(), a.cpu().numpy()) < 1e-6
    assert rel(Lb.cpu().numpy(), La.cpu().numpy()) < 1e-9


@pytest.mark.parametrize("prec", ["int8x5", "int8x6l", "fp16x5", "bf16x5"])
def test_zortho_step_split_matches_separate_calls(handle, prec):
    """The fused split step reuses psi's digits from its overlap in the first
    back-multiply and turns them into J psi's in place (pair swap, negation)
    for the second; the separate calls slice psi and J psi from FP64.  Same
    bits."""
    n, npw = 160, 3000
    psi = dev(make_kpoint_states(n, npw, eps=1e-3, seed=17).numpy())
    S, T, out, it, _ = handle.zortho_step(psi, nspin=2, tol=1e-13, maxit=50, modulus=True,
                                          prec=prec)
    S2 = handle.zoverlap(psi, prec=prec)
    T2, it2, _ = handle.zs_to_t(S2, nspin=2, tol=1e-13, maxit=50, modulus=True)
    out2 = handle.zbackmul(T2, psi, prec=prec)
    torch.cuda.synchronize()
    assert it == it2
    assert torch.equal(S, S2) and torch.equal(T, T2) and torch.equal(out, out2)


@pytest.mark.parametrize("prec", ["int8x5", "fp16x5"])
def test_zoverlap_split_imag_from_overlap_digits(handle, prec):
    """The split Hermitian overlap takes Im S from the digits of its own Re S
    slicing (J psi's made by a pair swap); the general product slices psi and
    J psi from FP64.  The imaginary parts agree bit for bit."""
    n, npw = 200, 2500
    psi = dev(make_kpoint_states(n, npw, eps=1e-3, seed=19).numpy())
    S = handle.zoverlap(psi, prec=prec)
    L = handle.zlambda(psi, psi, prec=prec)
    torch.cuda.synchronize()
    iu = torch.triu_indices(n, n, device=DEV)
    assert torch.equal(S.imag[iu[0], iu[1]], L.imag[iu[0], iu[1]])
    assert not S.imag.diagonal().any().item()


@pytest.mark.parametrize("prec", ["int8x5", "fp16x5"])
@pytest.mark.parametrize("transM", [False, True])
def test_zbackmul_split_equals_two_real_products(handle, prec, transM):
    """The split complex back-multiply slices psi once and makes J psi's
    digits from psi's in place; the real back-multiply of an explicit J psi
    slices J psi from FP64.  Re(op M) psi + Im(op M) J psi, same bits."""
    n, npw = 96, 2000
    rng = np.random.default_rng(21)
    psi_h = make_kpoint_states(n, npw, eps=0.0, seed=21).numpy()
    j_h = np.empty_like(psi_h)
    j_h[:, 0::2], j_h[:, 1::2] = -psi_h[:, 1::2], psi_h[:, 0::2]
    M_h = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    got = handle.zbackmul(dev(M_h), dev(psi_h), transM=transM, alpha=0.5, prec=prec)
    # op(M) = M^H: Re = Mr^T, Im = -Mi^T
    want = handle.backmul(dev(M_h.real), dev(psi_h), transM=transM, alpha=0.5, prec=prec)
    want = handle.backmul(dev(M_h.imag), dev(j_h), transM=transM, alpha=-0.5 if transM else 0.5,
                          beta=1.0, out=want, prec=prec)
    torch.cuda.synchronize()
    assert torch.equal(got, want)
