"""Pins the CPU oracle before anything is checked against it.

  * oao_dgemm is bitwise equal to the reference's compiled netlib dgemm_
    (src_mathlib/dgemm.f) on all four transpose cases, alpha/beta special
    cases and the XERBLA argument checks;
  * the overlap matches a long-double brute force of the 2Re - G0 metric, and
    its G=0 term reduces to the full-sphere complex overlap of a
    Hermitian-symmetric state set;
  * S->T matches the closed form V diag((lambda/2)^-1/2) V^T (what
    get_diag_Tmat computes, cp_lowdin.C:136-153), its invariants and special
    cases, PINY's serial iteration, and the reference's stop rule;
  * the whole oracle step run on oao_dgemm and on the reference dgemm_ is
    bitwise identical, and reproduces the committed golden fixtures.
"""
import glob
import os

import numpy as np
import pytest

from openatom_amd.workload import make_kpoint_states, make_states, metric_weights

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def need_ref(o):
    if not o.ref_available():
        pytest.skip("reference BLAS not built (oracle/_ref)")


# ------------------------------------------------------------------ DGEMM
@pytest.mark.parametrize("ta,tb", [("N", "N"), ("N", "T"), ("T", "N"), ("T", "T")])
@pytest.mark.parametrize("alpha,beta", [(1.0, 0.0), (2.0, 0.0), (-1.0, 1.0), (0.5, -0.25)])
def test_dgemm_bitwise_vs_reference(oracle, ta, tb, alpha, beta):
    need_ref(oracle)
    rng = np.random.default_rng(5)
    m, n, k = 23, 19, 31
    lda = (m if ta == "N" else k) + 3
    ldb = (k if tb == "N" else n) + 2
    ldc = m + 1
    A = rng.standard_normal(lda * (k if ta == "N" else m))
    B = rng.standard_normal(ldb * (n if tb == "N" else k))
    C0 = rng.standard_normal(ldc * n)
    C1, C2 = C0.copy(), C0.copy()
    oracle.dgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C1, ldc, which="oracle")
    oracle.dgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, C2, ldc, which="ref")
    assert np.array_equal(C1, C2)
    # and both equal the mathematical product to rounding
    Am = A.reshape(-1, lda).T[:, :] if True else None
    opA = (A.reshape(-1, lda).T[: (m if ta == "N" else k)])
    opA = opA if ta == "N" else opA.T
    opB = (B.reshape(-1, ldb).T[: (k if tb == "N" else n)])
    opB = opB if tb == "N" else opB.T
    Cm = C0.reshape(n, ldc).T[:m]
    ref = alpha * (opA @ opB) + beta * Cm
    got = C1.reshape(n, ldc).T[:m]
    assert np.allclose(got, ref, rtol=1e-12, atol=1e-12)
    # rows m..ldc-1 are not part of C and stay untouched
    assert np.array_equal(C1.reshape(n, ldc)[:, m:], C0.reshape(n, ldc)[:, m:])


def test_dgemm_argument_checks(oracle):
    need_ref(oracle)
    A = np.ones(16)
    B = np.ones(16)
    cases = [  # (ta, tb, m, n, k, lda, ldb, ldc) -> INFO (dgemm.f:169-190)
        (("X", "N", 2, 2, 2, 2, 2, 2), 1), (("N", "Q", 2, 2, 2, 2, 2, 2), 2),
        (("N", "N", -1, 2, 2, 2, 2, 2), 3), (("N", "N", 2, -1, 2, 2, 2, 2), 4),
        (("N", "N", 2, 2, -1, 2, 2, 2), 5), (("N", "N", 4, 2, 2, 3, 2, 4), 8),
        (("N", "N", 2, 2, 4, 2, 3, 2), 10), (("N", "N", 4, 2, 2, 4, 2, 3), 13),
        (("T", "N", 2, 2, 4, 3, 4, 2), 8), (("N", "T", 2, 4, 2, 2, 3, 2), 10),
    ]
    for (ta, tb, m, n, k, lda, ldb, ldc), info in cases:
        C = np.full(32, 7.0)
        got = oracle.dgemm_cm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, C, ldc, "oracle")
        assert got == info
        Cr = np.full(32, 7.0)
        oracle.dgemm_cm(ta, tb, m, n, k, 1.0, A, lda, B, ldb, 0.0, Cr, ldc, "ref")
        assert np.all(Cr == 7.0)  # XERBLA is silent: C untouched
        assert np.all(C == 7.0)


def test_dgemm_quick_returns(oracle):
    need_ref(oracle)
    A = np.full(9, np.nan)
    B = np.ones(9)
    for which in ("oracle", "ref"):
        C = np.arange(9.0)
        oracle.dgemm_cm("N", "N", 3, 3, 3, 0.0, A, 3, B, 3, 1.0, C, 3, which)
        assert np.array_equal(C, np.arange(9.0))  # alpha = 0, beta = 1: untouched
        oracle.dgemm_cm("N", "N", 3, 3, 3, 0.0, A, 3, B, 3, 2.0, C, 3, which)
        assert np.array_equal(C, 2 * np.arange(9.0))  # A never read
        oracle.dgemm_cm("N", "N", 3, 3, 3, 0.0, A, 3, B, 3, 0.0, C, 3, which)
        assert np.array_equal(C, np.zeros(9))


# ------------------------------------------------------------------ overlap
def _bruteforce_overlap(psi, nkx0):
    x = psi.astype(np.longdouble)
    w = metric_weights(psi.shape[1] // 2, nkx0).astype(np.longdouble)
    return (x * w) @ x.T


def test_overlap_bruteforce_longdouble(oracle):
    rng = np.random.default_rng(1)
    n, npw = 9, 70
    psi = rng.standard_normal((n, 2 * npw))
    for nkx0 in (0, 1, 5):
        S = oracle.overlap(psi, nkx0=nkx0)
        ref = np.array(_bruteforce_overlap(psi, nkx0), dtype=np.float64)
        assert np.allclose(S, ref, rtol=0, atol=1e-13 * np.abs(ref).max())
        assert np.array_equal(S, S.T)  # exactly symmetric
    # a dropped kx0 correction is caught
    assert not np.allclose(oracle.overlap(psi, nkx0=0), _bruteforce_overlap(psi, 1).astype(float),
                           atol=1e-6)


def test_overlap_plane_split_invariant(oracle):
    rng = np.random.default_rng(2)
    psi = rng.standard_normal((12, 2 * 257))
    S1 = oracle.overlap(psi, 1, nplanes=1)
    S7 = oracle.overlap(psi, 1, nplanes=7)
    assert np.allclose(S1, S7, rtol=0, atol=1e-13 * np.abs(S1).max())


def test_overlap_g0_term_equals_full_sphere(oracle):
    """Half sphere {G=0} U {+G} with the 2Re - G0 weight equals the complex
    overlap of the Hermitian-symmetric full sphere (psi(-G) = psi(G)^*)."""
    rng = np.random.default_rng(3)
    n, npw = 6, 40
    half = rng.standard_normal((n, npw)) + 1j * rng.standard_normal((n, npw))
    half[:, 0] = half[:, 0].real  # G = 0 real
    full = np.concatenate([half, np.conj(half[:, 1:])], axis=1)  # G=0, +G, -G
    S_full = (np.conj(full) @ full.T)
    assert np.allclose(S_full.imag, 0, atol=1e-12)
    psi = np.empty((n, 2 * npw))
    psi[:, 0::2], psi[:, 1::2] = half.real, half.imag
    S = oracle.overlap(psi, nkx0=1)
    assert np.allclose(S, S_full.real, rtol=1e-13, atol=1e-12)


def test_overlap_orthonormal_block_is_identity(oracle):
    psi = make_states(24, 400, eps=0.0, seed=4).numpy()
    S = oracle.overlap(psi, nkx0=1)
    assert np.abs(S - np.eye(24)).max() < 1e-13


# ------------------------------------------------------------------ S -> T
def _closed_form_T(S, nspin):
    B = S / 2.0 if nspin == 1 else S
    lam, V = np.linalg.eigh(B)
    return (V * lam ** -0.5) @ V.T


@pytest.mark.parametrize("nspin", [1, 2])
def test_s_to_t_closed_form(oracle, nspin):
    psi = make_states(48, 600, eps=1e-2, seed=6).numpy()
    S = oracle.overlap(psi, 1)
    if nspin == 1:
        S = 2.0 * S  # occupation-2 convention: S ~ 2I
    T, it, err, st = oracle.s_to_t(S, nspin=nspin, tol=1e-14, maxit=100)
    assert st == 0 and err <= 1e-14
    Tc = _closed_form_T(S, nspin)
    assert np.abs(T - Tc).max() < 1e-13
    B = S / 2.0 if nspin == 1 else S
    assert np.abs(T @ B @ T - np.eye(48)).max() < 1e-13  # T B T = I
    assert np.abs(T - T.T).max() < 1e-14
    assert np.abs(T @ S - S @ T).max() < 1e-13


def test_s_to_t_special_cases(oracle):
    n = 16
    T, it, err, st = oracle.s_to_t(2.0 * np.eye(n), nspin=1, tol=1e-15, maxit=50)
    assert st == 0 and it == 1 and err == 0.0 and np.array_equal(T, np.eye(n))
    d = 1.0 + 0.05 * np.linspace(-1, 1, n)
    T, it, err, st = oracle.s_to_t(2.0 * np.diag(d), nspin=1, tol=1e-15, maxit=50)
    assert st == 0
    assert np.allclose(np.diag(T), d ** -0.5, rtol=0, atol=1e-15)
    assert np.abs(T - np.diag(np.diag(T))).max() == 0.0
    # lsda: B = S
    T, it, err, st = oracle.s_to_t(np.diag(d), nspin=2, tol=1e-15, maxit=50)
    assert np.allclose(np.diag(T), d ** -0.5, rtol=0, atol=1e-15)


def test_s_to_t_stop_rule(oracle):
    """ortho.C:118-133: continue while err > tol and it < maxit; abort status
    whenever the count reached maxit at exit, even on a converged iterate."""
    psi = make_states(20, 300, eps=1e-2, seed=8).numpy()
    S = oracle.overlap(psi, 1)
    T, it, err, st = oracle.s_to_t(S, nspin=2, tol=1e-15, maxit=100)
    assert st == 0 and it < 100
    T2, it2, err2, st2 = oracle.s_to_t(S, nspin=2, tol=1e-15, maxit=it)
    assert it2 == it and st2 == 1 and np.array_equal(T, T2)
    T3, it3, err3, st3 = oracle.s_to_t(S, nspin=2, tol=1e-15, maxit=2)
    assert it3 == 2 and st3 == 1 and err3 > 1e-15


def test_piny_iteration_matches_ortho(oracle):
    psi = make_states(70, 500, eps=1e-3, seed=9).numpy()
    S = 2.0 * oracle.overlap(psi, 1)
    Tp, itp, tolp, stp = oracle.piny_iter_tmat(S)
    To, ito, erro, sto = oracle.s_to_t(S, nspin=1, tol=1e-15, maxit=10)
    assert stp == 0
    assert np.abs(Tp - To).max() < 1e-14
    assert abs(itp - ito) <= 1


# ------------------------------------------------------------------ backward
def test_backmul_linear_map(oracle):
    rng = np.random.default_rng(10)
    n, K = 11, 2 * 97
    M = rng.standard_normal((n, n))
    psi = rng.standard_normal((n, K))
    out0 = rng.standard_normal((n, K))
    out = oracle.backmul(M, psi, transT=False, alpha=0.5, beta=-1.0, out=out0.copy())
    assert np.allclose(out, 0.5 * M @ psi - out0, rtol=1e-13, atol=1e-13)
    out = oracle.backmul(M, psi, transT=True, alpha=1.0, beta=0.0, nplanes=3)
    assert np.allclose(out, M.T @ psi, rtol=1e-13, atol=1e-13)
    assert not np.allclose(out, M @ psi, atol=1e-3)  # a transposed operand is caught


# ------------------------------------------------------------------ whole step
def test_step_oracle_equals_reference_bitwise(oracle):
    need_ref(oracle)
    psi = make_states(33, 257, eps=1e-3, seed=14).numpy()
    a = oracle.ortho_step(psi, nkx0=1, nspin=2, gemm="oracle")
    b = oracle.ortho_step(psi, nkx0=1, nspin=2, gemm="ref")
    for k in ("S", "T", "psi_out"):
        assert np.array_equal(a[k], b[k])
    assert a["iters"] == b["iters"]


def test_config_a_step_orthonormalises(oracle):
    """Config A (n=128, npw=8192, eps=1e-3, seed 0): one step leaves
    ||S(T psi) - I||_max <= 1e-12 (lsda convention, S ~ I)."""
    psi = make_states(128, 8192, eps=1e-3, seed=0).numpy()
    r = oracle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=1000, nplanes=8)
    assert r["status"] == 0
    res0 = oracle.orthonormality_residual(psi, 1, occ=1.0, nplanes=8)
    res1 = oracle.orthonormality_residual(r["psi_out"], 1, occ=1.0, nplanes=8)
    assert 1e-4 < res0 < 1e-2
    assert res1 < 1e-12


@pytest.mark.parametrize("path", sorted(glob.glob(os.path.join(GOLDEN, "step_*.npz"))))
def test_golden_fixtures(oracle, path):
    g = np.load(path)
    psi = g["psi"]
    nkx0 = int(g["nkx0"])
    for nspin in (1, 2):
        r = oracle.ortho_step(psi, nkx0=nkx0, nspin=nspin, gemm="oracle")
        assert np.array_equal(r["S"], g[f"S_nspin{nspin}"])
        assert np.array_equal(r["T"], g[f"T_nspin{nspin}"])
        assert np.array_equal(r["psi_out"], g[f"psi_out_nspin{nspin}"])
        assert r["iters"] == int(g[f"iters_nspin{nspin}"])
    L = oracle.lambda_fwd(psi, g["F"])
    assert np.array_equal(L, g["L"])
    assert np.array_equal(oracle.backmul(L, psi, transT=True), g["LT_psi"])


# ---------------------------------------------------------------- ZGEMM
ZCASES = [(1 + 0j, 0j), (0.5 - 1.5j, 0.25 + 2j), (0j, 1.5j), (0j, 0j), (1 + 0j, 1 + 0j),
          (2j, 0j)]


@pytest.mark.parametrize("ta", list("NTC"))
@pytest.mark.parametrize("tb", list("NTC"))
def test_zgemm_oracle_matches_reference_bitwise(oracle, ta, tb):
    """oao_zgemm restates zgemm.f's nine loop nests; the compiled reference
    zgemm_ must agree to the last bit (exact zero entries of B take the
    reference's skip branch)."""
    if not oracle.ref_available():
        pytest.skip("reference BLAS not built")
    rng = np.random.default_rng(ord(ta) * 3 + ord(tb))
    m, n, k = 7, 5, 9
    ra, ca = (m, k) if ta == "N" else (k, m)
    rb, cb = (k, n) if tb == "N" else (n, k)
    lda, ldb, ldc = ra + 2, rb + 1, m + 3
    for alpha, beta in ZCASES:
        A = rng.standard_normal((ca, lda)) + 1j * rng.standard_normal((ca, lda))
        B = rng.standard_normal((cb, ldb)) + 1j * rng.standard_normal((cb, ldb))
        B[0, 0] = 0
        C0 = rng.standard_normal((n, ldc)) + 1j * rng.standard_normal((n, ldc))
        c1, c2 = C0.copy(), C0.copy()
        assert oracle.zgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, c1, ldc) == 0
        oracle.zgemm_cm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, c2, ldc, which="ref")
        assert np.array_equal(c1.view(np.float64), c2.view(np.float64)), (alpha, beta)


def test_zgemm_oracle_is_the_complex_product(oracle):
    """Against numpy's complex matmul (independent arithmetic): a wrong
    conjugation or transposition shows up as an O(1) error.  A column-major
    matrix with r rows and c columns is held as a C-contiguous (c, r) array."""
    rng = np.random.default_rng(5)
    m, n, k = 6, 4, 11

    def cn(*shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    for ta in "NTC":
        for tb in "NTC":
            if ta == "N":
                As, lda = cn(k, m), m
                opA = As.T
            else:
                As, lda = cn(m, k), k
                opA = As if ta == "T" else As.conj()
            if tb == "N":
                Bs, ldb = cn(n, k), k
                opB = Bs.T
            else:
                Bs, ldb = cn(k, n), n
                opB = Bs if tb == "T" else Bs.conj()
            Cs = np.zeros((n, m), dtype=np.complex128)
            oracle.zgemm_cm(ta, tb, m, n, k, 1.0, As, lda, Bs, ldb, 0.0, Cs, m)
            assert np.abs(Cs.T - opA @ opB).max() < 1e-12, (ta, tb)


def test_zgemm_argument_checks(oracle):
    z = np.zeros(16, dtype=np.complex128)
    assert oracle.zgemm_cm("X", "N", 2, 2, 2, 1, z, 2, z, 2, 0, z, 2) == 1
    assert oracle.zgemm_cm("N", "Q", 2, 2, 2, 1, z, 2, z, 2, 0, z, 2) == 2
    assert oracle.zgemm_cm("N", "N", -1, 2, 2, 1, z, 2, z, 2, 0, z, 2) == 3
    assert oracle.zgemm_cm("C", "N", 2, 2, 3, 1, z, 2, z, 3, 0, z, 2) == 8
    assert oracle.zgemm_cm("N", "C", 2, 3, 2, 1, z, 2, z, 2, 0, z, 2) == 10
    assert oracle.zgemm_cm("N", "N", 3, 2, 2, 1, z, 3, z, 2, 0, z, 2) == 13


@pytest.mark.parametrize("kpt", [0, 1])
def test_gwbse_states_overlap(oracle, kpt):
    """The reference's own complex state files (gw_bse input, 8 states x 718
    plane waves per k-point): the oracle's ZGEMM('C','N') reproduces the
    reference ZGEMM's S bit for bit, and the states are orthonormal."""
    d = np.load(os.path.join(GOLDEN, f"gwbse_kpt{kpt}.npz"))
    psi = np.ascontiguousarray(d["psi"])
    nst, ncoef = psi.shape
    S = np.zeros((nst, nst), dtype=np.complex128)
    oracle.zgemm_cm("C", "N", nst, nst, ncoef, 1.0, psi, ncoef, psi, ncoef, 0.0, S, nst)
    assert np.array_equal(S.view(np.float64), d["S_cm"].view(np.float64))
    assert np.abs(S.T - np.eye(nst)).max() < 5e-15
    assert np.abs(S - S.conj().T).max() < 1e-16  # Hermitian


# ---- GSpace accumulation of the Lambda round trip --------------------------
#
# The pin is the constraint the accumulation enforces.  Lambda is the plain
# dot product L[x][y] = sum_d psi_x[d] F_y[d] over the stored doubles, and the
# stored force carries the two members of a +-G pair in one kx > 0 point.  For
# states orthonormal in the path metric (S = 2I under lda, S = I under lsda,
# Lambda doubled by Ortho under lsda) the weights 1 (kx > 0) and 1/2 (kx = 0)
# make the updated force Lambda-free: L(psi, F') = 0.  A wrong weight, a
# dropped kx = 0 branch or a transposed Lambda leaves O(|L|) behind.

def _gsp_states(n, npw, nkx0, double_pack, lsda, seed):
    # single pack: every point has weight 1 in the metric (no +-G pairing)
    psi = make_states(n, npw, eps=0.0, seed=seed, nkx0=nkx0 if double_pack else npw).numpy()
    return psi * (1.0 if lsda else np.sqrt(2.0))


@pytest.mark.parametrize("mode,lsda,double_pack", [
    (0, False, True), (0, True, True), (0, False, False), (0, True, False),
    (1, False, True), (1, True, True), (1, False, False), (1, True, False)])
def test_gsp_unpack_minimisation_projects_out_lambda(oracle, mode, lsda, double_pack):
    n, npw, nkx0 = 10, 257, 7
    psi = _gsp_states(n, npw, nkx0, double_pack, lsda, seed=5)
    F = np.random.default_rng(6).standard_normal((n, 2 * npw))
    L = oracle.lambda_fwd(psi, F)
    if mode == 0 and lsda:
        L = 2.0 * L  # Ortho::acceptSectionLambda doubles Lambda under lsda
    data = oracle.backmul(L, psi, transT=True)
    F1 = F.copy()
    oracle.gsp_unpack(data, F1, nkx0=nkx0, double_pack=double_pack, mode=mode, lsda=lsda)
    L1 = oracle.lambda_fwd(psi, F1)
    assert np.abs(L1).max() < 1e-12 * np.abs(L).max()
    # the wrong kx = 0 weight leaves the kx = 0 share of Lambda behind
    if double_pack:
        bad = F.copy()
        oracle.gsp_unpack(data, bad, nkx0=0, double_pack=double_pack, mode=mode, lsda=lsda)
        assert np.abs(oracle.lambda_fwd(psi, bad)).max() > 1e-6 * np.abs(L).max()


@pytest.mark.parametrize("lsda", [False, True])
def test_gsp_unpack_dynamics_projects_out_lambda(oracle, lsda):
    """Dynamics: psi and F are sent with the kx = 0 points scaled by 1/sqrt2
    and sqrt2; the PairCalculator returns data = -T F + Gamma^T psi with
    Gamma = T Lambda; force = -data, then the kx = 0 points times 1/sqrt2.
    For orthonormal states T = I and the force is again Lambda-free."""
    n, npw, nkx0 = 9, 200, 6
    psi = _gsp_states(n, npw, nkx0, True, lsda, seed=8)
    F = np.random.default_rng(9).standard_normal((n, 2 * npw))
    ps, Fs = psi.copy(), F.copy()
    ps[:, :2 * nkx0] *= 1.0 / np.sqrt(2.0)
    Fs[:, :2 * nkx0] *= np.sqrt(2.0)
    S = oracle.overlap(psi, nkx0=nkx0)
    T, _, _, st = oracle.s_to_t(S, nspin=2 if lsda else 1)
    assert st == 0 and np.abs(T - np.eye(n)).max() < 1e-13
    L = oracle.lambda_fwd(ps, Fs)
    G = T @ (2.0 * L if lsda else L)
    data = oracle.backmul(T, Fs, transT=False, alpha=-1.0)
    oracle.backmul(G, ps, transT=True, alpha=1.0, beta=1.0, out=data)
    force = np.zeros_like(F)
    oracle.gsp_unpack(data, force, nkx0=nkx0, double_pack=True, mode=2, lsda=lsda)
    assert np.abs(oracle.lambda_fwd(psi, force)).max() < 1e-12 * np.abs(L).max()


# ---- complex-math (k-point) build ----------------------------------------

def _cplx(psi):
    return psi[:, 0::2] + 1j * psi[:, 1::2]


def test_zoverlap_is_psi_psih(oracle):
    psi = make_kpoint_states(20, 333, eps=1e-2, seed=2).numpy()
    S = oracle.zoverlap(psi)
    c = _cplx(psi)
    want = c @ c.conj().T
    assert np.abs(S - want).max() < 1e-14 * np.abs(want).max()
    assert np.array_equal(S, S.conj().T)  # ZGEMM('C','N') terms are exact conjugates
    L = oracle.zlambda(psi, psi[::-1].copy())
    want = c @ c[::-1].conj().T
    assert np.abs(L - want).max() < 1e-14 * np.abs(want).max()
    if oracle.ref_available():
        assert np.array_equal(oracle.zoverlap(psi, gemm="ref"), S)


@pytest.mark.parametrize("kpt", [0, 1])
def test_zoverlap_gwbse_fixture(oracle, kpt):
    """The reference's k-point states: the forward of the complex-math build
    gives the identity."""
    d = np.load(os.path.join(GOLDEN, f"gwbse_kpt{kpt}.npz"))
    psi = np.ascontiguousarray(d["psi"]).view(np.float64)
    S = oracle.zoverlap(psi)
    assert np.abs(S - np.eye(S.shape[0])).max() < 5e-15
    res = oracle.zortho_step(psi, nspin=2, tol=1e-15, maxit=20)
    assert res["status"] == 0
    assert np.abs(res["T"] - np.eye(S.shape[0])).max() < 1e-14


@pytest.mark.parametrize("nspin", [1, 2])
def test_zs_to_t_closed_form(oracle, nspin):
    """T = B^(-1/2) of the Hermitian B = S/2 (nspin 1) or S, from numpy's
    Hermitian eigensolver."""
    psi = make_kpoint_states(40, 500, eps=2e-2, seed=3).numpy()
    S = oracle.zoverlap(psi)
    if nspin == 1:
        S = 2.0 * S
    T, it, err, st = oracle.zs_to_t(S, nspin=nspin, tol=1e-14, maxit=100, modulus=True)
    assert st == 0 and err <= 1e-14
    B = S / 2.0 if nspin == 1 else S
    lam, V = np.linalg.eigh(B)
    Tc = (V * lam ** -0.5) @ V.conj().T
    assert np.abs(T - Tc).max() < 1e-13
    assert np.abs(T @ B @ T - np.eye(40)).max() < 1e-13
    assert np.abs(T - T.conj().T).max() < 1e-14


def test_zs_to_t_real_input_equals_real_oracle(oracle):
    """On a real symmetric S the complex loop performs the real one's
    arithmetic exactly (imaginary parts stay zero, the complex products add
    exact zeros), so it must reproduce the real oracle bit for bit -- which is
    itself pinned to the compiled reference."""
    psi = make_states(32, 400, eps=1e-2, seed=4).numpy()
    S = oracle.overlap(psi, 1)
    Tr, itr, errr, str_ = oracle.s_to_t(S, nspin=2, tol=1e-15, maxit=50)
    Tz, itz, errz, stz = oracle.zs_to_t(S.astype(np.complex128), nspin=2, tol=1e-15, maxit=50)
    assert (itr, str_) == (itz, stz) and errr == errz
    assert np.array_equal(Tz.real, Tr) and not Tz.imag.any()
    if oracle.ref_available():
        Tf, itf, errf, stf = oracle.zs_to_t(S.astype(np.complex128), nspin=2, tol=1e-15,
                                            maxit=50, gemm="ref")
        assert np.array_equal(Tf, Tz) and itf == itz


def test_zs_to_t_reference_stop_rule(oracle):
    """The reference's complex rule takes the real part of sum (dA)^2: it ends
    the loop either below the tolerance or on a NaN error (negative sum);
    the modulus rule always runs to the tolerance."""
    psi = make_kpoint_states(24, 300, eps=5e-2, seed=5).numpy()
    S = oracle.zoverlap(psi)
    Tm, itm, errm, stm = oracle.zs_to_t(S, nspin=2, tol=1e-14, maxit=60, modulus=True)
    Tr, itr, errr, str_ = oracle.zs_to_t(S, nspin=2, tol=1e-14, maxit=60)
    assert stm == 0 and errm <= 1e-14
    assert str_ == 0 and (errr <= 1e-14 or np.isnan(errr)) and itr <= itm


@pytest.mark.parametrize("transM", [False, True])
def test_zbackmul_is_the_complex_product(oracle, transM):
    rng = np.random.default_rng(7)
    psi = make_kpoint_states(16, 200, eps=0.0, seed=6).numpy()
    M = rng.standard_normal((16, 16)) + 1j * rng.standard_normal((16, 16))
    out0 = rng.standard_normal(psi.shape)
    got = oracle.zbackmul(M, psi, transM=transM, alpha=-0.5, beta=2.0, out=out0.copy())
    op = M.conj().T if transM else M
    want = -0.5 * (op @ _cplx(psi)) + 2.0 * _cplx(out0)
    assert np.abs(_cplx(got) - want).max() < 1e-13 * np.abs(want).max()


@pytest.mark.parametrize("nspin", [1, 2])
def test_zortho_step_orthonormalises(oracle, nspin):
    psi = make_kpoint_states(30, 400, eps=1e-2, seed=8).numpy()
    res = oracle.zortho_step(psi, nspin=nspin, tol=1e-15, maxit=50, modulus=True)
    assert res["status"] == 0
    c = _cplx(res["psi_out"])
    occ = 2.0 if nspin == 1 else 1.0
    assert np.abs(c @ c.conj().T - occ * np.eye(30)).max() < 1e-13


def test_zortho_step_reference_rule_under_lsda(oracle):
    """Under lsda B = S, and the first change dA = (I - S)/2 is dominated by
    off-diagonal complex entries, whose squares have negative real parts:
    the reference's complex rule reads a NaN error and stops after one
    iteration, leaving the block far from orthonormal."""
    psi = make_kpoint_states(30, 400, eps=1e-2, seed=8).numpy()
    res = oracle.zortho_step(psi, nspin=2, tol=1e-15, maxit=50)
    assert res["status"] == 0 and res["iters"] == 1 and np.isnan(res["err"])


# ------------------------------------------------------------ array_diag_max
def test_array_diag_max_pins(oracle):
    """Ortho::array_diag_max (ortho.h:282-313) reduced over the Ortho tiles:
    - for S with a non-negative diagonal it is max|S - d I| (numpy), for every
      tile size og (the max reduction makes the tiling invisible);
    - d = 1 under lsda, 2 otherwise;
    - on a diagonal tile the diagonal term is ||S_ii| - d|, so a negative
      diagonal entry -x counts as |x - d|, not x + d; the tile's first element
      also enters as |S_00 - d|, which only matters when S_00 < 0;
    - complex S takes the modulus (myabs = std::abs)."""
    rng = np.random.default_rng(3)
    n = 11
    X = rng.standard_normal((n, n)) * 1e-3
    S = 2.0 * np.eye(n) + (X + X.T)
    want = np.abs(S - 2.0 * np.eye(n)).max()
    for og in (0, 1, 3, 4, 11):
        assert oracle.array_diag_max(S, og=og) == want
    assert oracle.array_diag_max(S, lsda=True) == np.abs(S - np.eye(n)).max()
    # a negative diagonal entry in a tile that does not start with it
    S2 = S.copy()
    S2[5, 5] = -2.5
    got = oracle.array_diag_max(S2, og=0)
    assert got == max(np.abs(S2 - 2.0 * np.eye(n))[np.arange(n) != 5].max(),
                      np.abs(S2[5, np.arange(n) != 5]).max(), abs(2.5 - 2.0))
    # ... and as the first element of a tile: |S_00 - d| = 4.5 enters too
    S3 = S.copy()
    S3[0, 0] = -2.5
    assert oracle.array_diag_max(S3, og=0) == 4.5
    # complex: modulus
    Y = (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) * 1e-3
    Z = np.eye(n) + (Y + Y.conj().T)
    zwant = np.abs(Z - np.eye(n)).max()
    assert abs(oracle.array_diag_max(Z, lsda=True) - zwant) <= 1e-18
