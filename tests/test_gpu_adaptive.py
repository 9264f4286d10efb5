"""Adaptive per-chunk slice counts (OA_PREC_ADAPTIVE) of the int8 split.

The rule: with W_z = K_z 4^(e_z) (K_z columns of chunk z, e_z its exponent)
and W = sum_z W_z, chunk z drops d slices while W_z 254^d <= (254/4) W / nz,
keeping at least 2.  The element-wise truncation bound is then at most 64.5
times that of s uniform slices, about a quarter of the bound of s - 1
uniform slices, so int8x6a is checked against the oracle (or the FP64 path,
itself pinned to the oracle at 1e-12) with the uniform int8x5 tolerance; a
block without the |G| envelope must keep every slice and give the uniform
split's bits.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"
RADIX = 254.0
THETA = RADIX / 4.0


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


def chunk_exp(m: float) -> int:
    """Exponent e with |x| 2^(7-e) <= 127 for |x| <= m (0 for m = 0)."""
    if not m > 0.0:
        return 0
    f, E = np.frexp(m)
    return int(E) if f <= 127.0 / 128.0 else int(E) + 1


def host_counts(psi: np.ndarray, kchunk: int, smax: int) -> list[int]:
    """The count rule restated on the host from the block's chunk maxima."""
    K = psi.shape[1]
    nz = len(range(0, ((K + 127) // 128) * 128, kchunk))
    maxes = [float(np.abs(psi[:, z * kchunk:min(K, (z + 1) * kchunk)]).max())
             if z * kchunk < K else 0.0 for z in range(nz)]
    es = [chunk_exp(m) for m in maxes if m > 0.0]
    eref = max(es)
    w = [np.ldexp(float(max(0, min(K, (z + 1) * kchunk) - z * kchunk)),
                  2 * (chunk_exp(m) - eref)) if m > 0.0 else 0.0
         for z, m in enumerate(maxes)]
    budget = THETA * sum(w) / nz
    out = []
    for wz in w:
        s = smax
        while s > 2 and wz * RADIX <= budget:
            wz *= RADIX
            s -= 1
        out.append(s)
    return out


@pytest.fixture(scope="module")
def block():
    return make_states(256, 16384, eps=1e-3, seed=3, device=DEV, rng="torch")


def flat_block(n, npw, seed):
    """Orthonormal block with no |G| envelope: every chunk has the same scale."""
    g = torch.Generator(device=DEV)
    g.manual_seed(seed)
    Z = torch.randn((n, 2 * npw), generator=g, dtype=torch.float64, device=DEV)
    Z[:, 1] = 0.0
    w = torch.full((2 * npw,), 2.0, dtype=torch.float64, device=DEV)
    w[:2] = 1.0
    for _ in range(2):
        L = torch.linalg.cholesky((Z * w) @ Z.T)
        Z = torch.linalg.solve_triangular(L, Z, upper=False)
    return Z.contiguous()


@pytest.mark.parametrize("smax", [4, 5, 6])
def test_counts_follow_the_rule(handle, block, smax):
    counts, kchunk = handle.slice_counts(block, nkx0=1, prec=f"int8x{smax}a")
    want = host_counts(block.cpu().numpy(), kchunk, smax)
    print(f"int8x{smax}a counts {counts}")
    assert counts == want
    # the envelope makes the leading chunk the heaviest: it keeps every slice
    assert counts[0] == smax and min(counts) < smax


def test_flat_spectrum_keeps_every_slice(handle):
    psi = flat_block(256, 16384, seed=11)
    counts, _ = handle.slice_counts(psi, nkx0=1, prec="int8x6a")
    assert set(counts) == {6}
    a = handle.overlap_sym(psi, nkx0=1, prec="int8x6")
    b = handle.overlap_sym(psi, nkx0=1, prec="int8x6a")
    X = 1e-3 * torch.randn(256, 256, dtype=torch.float64, device=DEV)
    T = torch.eye(256, dtype=torch.float64, device=DEV) + (X + X.T)
    c = handle.backmul(T, psi, prec="int8x6")
    d = handle.backmul(T, psi, prec="int8x6a")
    torch.cuda.synchronize()
    assert torch.equal(a, b)
    assert torch.equal(c, d)


def test_adaptive_overlap_and_backmul_vs_oracle(handle, oracle):
    """n=128, npw=8192 (the CPU reference case): S and T*psi of the adaptive
    split against the oracle's FP64 overlap and back-multiply (the netlib DGEMM
    loop order), within the uniform int8x5 tolerance; the uniform split
    is checked against the oracle the same way."""
    n, npw = 128, 8192
    psi_h = make_states(n, npw, eps=1e-3, seed=0)
    psi = psi_h.to(DEV)
    S_ref = oracle.overlap(psi_h.numpy(), nkx0=1)
    T_ref = oracle.s_to_t(S_ref, nspin=2, tol=1e-13, maxit=50)[0]
    out_ref = oracle.backmul(T_ref, psi_h.numpy())
    T = torch.from_numpy(T_ref).to(DEV)
    for prec, tol_s, tol_o in (("int8x5", 3e-10, 3e-11), ("int8x6a", 3e-10, 3e-11)):
        S = handle.overlap_sym(psi, nkx0=1, prec=prec).cpu().numpy()
        out = handle.backmul(T, psi, prec=prec).cpu().numpy()
        es = float(np.abs(S - S_ref).max() / np.abs(S_ref).max())
        eo = float(np.abs(out - out_ref).max() / np.abs(out_ref).max())
        print(f"{prec}: S vs oracle {es:.3e}, T*psi vs oracle {eo:.3e}")
        assert es < tol_s and eo < tol_o
        assert np.array_equal(S, S.T)


def test_adaptive_step_config_b(handle):
    """n=512, npw=65536: residual <= 1e-10 and T within 1e-10 of FP64."""
    n, npw = 512, 65536
    psi = make_states(n, npw, eps=1e-3, seed=0, device=DEV, rng="torch")
    _, T64, _, _, _ = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=50)
    counts, _ = handle.slice_counts(psi, nkx0=1, prec="int8x6a")
    _, Ta, oa, _, _ = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=50,
                                        prec="int8x6a")
    eye = torch.eye(n, dtype=torch.float64, device=DEV)
    res = (handle.overlap_sym(oa, nkx0=1, prec="fp64") - eye).abs().max().item()
    dT = (Ta - T64).abs().max().item()
    print(f"int8x6a: counts histogram {np.bincount(counts).tolist()}, residual {res:.3e}, "
          f"|dT| {dT:.3e}")
    assert res <= 1e-10 and dT <= 1e-10


def test_adaptive_batch_equals_single(handle):
    """A grouped batch in the adaptive mode gives each instance's single-call
    bits (graph replay included)."""
    from openatom_amd.ortho import Batch
    psis = [make_states(n, p, eps=e, seed=i, device=DEV, rng="torch")
            for i, (n, p, e) in enumerate([(96, 8192, 1e-3), (256, 20480, 1e-2),
                                           (160, 65536, 1e-4)])]
    b = Batch(handle)
    b.plan(psis, nkx0=1, nspin=2, tol=1e-12, maxit=12, prec="int8x6a")
    b.launch()
    b.launch()
    torch.cuda.synchronize()
    for i, p in enumerate(psis):
        S, T, out, _, _ = handle.ortho_step(p, nkx0=1, nspin=2, tol=1e-12, maxit=12,
                                            prec="int8x6a")
        torch.cuda.synchronize()
        assert torch.equal(out, b.outs[i]), i
    b.close()


def test_reuse_hint_reslices_after_other_work(handle, block):
    """A back-multiply asked to reuse digits after the workspace was used for
    another block slices its own psi again (same bits as without reuse)."""
    other = make_states(256, 16384, eps=1e-3, seed=4, device=DEV, rng="torch")
    S = handle.overlap_sym(block, nkx0=1, prec="int8x6a")
    T, _, _ = handle.s_to_t(S, nspin=2, tol=1e-13, maxit=50)
    want = handle.backmul(T, block, prec="int8x6a")
    handle.overlap_sym(block, nkx0=1, prec="int8x6a")
    handle.overlap_sym(other, nkx0=1, prec="int8x6a")
    got = handle.backmul(T, block, prec="int8x6a", reuse_digits=True)
    handle.overlap_sym(block, nkx0=1, prec="int8x6a")
    handle.lambda_fwd(block, other, prec="int8x5")  # overwrites the digit area
    got2 = handle.backmul(T, block, prec="int8x6a", reuse_digits=True)
    torch.cuda.synchronize()
    assert torch.equal(got, want) and torch.equal(got2, want)


def test_adaptive_rejected_where_unsupported(handle, block):
    from openatom_amd._lib import OaError
    with pytest.raises(OaError):
        handle.overlap_sym(block, nkx0=1, prec="fp16x5a")
    with pytest.raises(OaError):
        handle.overlap_sym(block, nkx0=1, prec="int8x3a")
    with pytest.raises(OaError):
        handle.lambda_fwd(block, block, prec="int8x6a")


def test_backmul_int32_capacity(handle):
    """The int8 back-multiply's int32 level sums over n states: n = 40,000 at
    int8x5 (4 levels: 40,064 * 4 * 127^2 >= 2^31) is refused before any
    device work (the buffers passed are never touched)."""
    from openatom_amd import _lib
    L = _lib.lib()
    small = torch.zeros(16, dtype=torch.float64, device=DEV)
    p = small.data_ptr()
    st = L.oa_ortho_step(handle.h, 40000, 64, p, 128, 1, 2, 1e-12, 12,
                         _lib.prec_code("int8x5"), p, p, p, 128, None, None)
    assert st == _lib.OA_ERR_CAPACITY
    st = L.oa_backmul(handle.h, 40000, 64, p, 0, p, 128, 1.0, 0.0, p, 128,
                      _lib.prec_code("int8x5"))
    assert st == _lib.OA_ERR_CAPACITY
