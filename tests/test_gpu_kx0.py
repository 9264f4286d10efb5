"""A wide kx = 0 plane on the split path.

The leading 2*nkx0 columns count half in S.  The split path carries that
weight inside the digit products (the S-layout digits hold those columns
times 1/sqrt(2)), so its cost does not grow with nkx0.  These cases put the
plane across several chunks (nkx0 = 800 of npw = 3000: 1600 columns against
chunks of 512 columns at n = 96) and check S against the oracle's overlap,
the whole step against the oracle's step, and grouped launches against
single-instance ones bit for bit.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"
# |dS| / max|S| bounds of the digit arithmetic (see test_gpu_split.py)
TOL = {"int8x4": 3e-7, "int8x5": 3e-10, "int8x6a": 3e-10, "fp16x5": 3e-10}


def relerr(got, want):
    return float(np.abs(got - want).max() / np.abs(want).max())


@pytest.fixture(scope="module")
def wide():
    return make_states(96, 3000, eps=1e-3, seed=11, nkx0=800)


@pytest.mark.parametrize("prec", ["int8x4", "int8x5", "int8x6a", "fp16x5"])
def test_overlap_with_wide_kx0_plane_against_the_oracle(handle, oracle, wide, prec):
    want = oracle.overlap(wide.numpy(), nkx0=800)
    S = handle.overlap_sym(wide.to(DEV), nkx0=800, alpha=2.0, prec=prec)
    torch.cuda.synchronize()
    got = S.cpu().numpy()
    assert np.array_equal(got, got.T)
    e = relerr(got, want)
    print(f"{prec}: |S - oracle| / max|S| = {e:.2e}")
    assert e < TOL[prec]
    # the plane's half weight is in the result: counting it fully is far off
    full = oracle.overlap(wide.numpy(), nkx0=0)
    assert relerr(got, full) > 1e3 * TOL[prec]


def test_step_with_wide_kx0_plane_against_the_oracle(handle, oracle, wide):
    ref = oracle.ortho_step(wide.numpy(), nkx0=800, nspin=2, tol=1e-12, maxit=30)
    assert ref["status"] == 0
    S, T, out, it, err = handle.ortho_step(wide.to(DEV), nkx0=800, nspin=2, tol=1e-12, maxit=30,
                                           prec="int8x6a")
    torch.cuda.synchronize()
    for name, got, want in (("S", S, ref["S"]), ("T", T, ref["T"]), ("T psi", out, ref["psi_out"])):
        e = relerr(got.cpu().numpy(), want)
        print(f"{name}: {e:.2e}")
        assert e < 3e-10, name
    # orthonormal in the path's metric, which weights the plane by 1/2
    r = (handle.overlap_sym(out, nkx0=800, alpha=2.0, prec="fp64")
         - torch.eye(96, dtype=torch.float64, device=DEV)).abs().max().item()
    assert r <= 1e-10


@pytest.mark.parametrize("prec", ["int8x5", "int8x6a"])
def test_grouped_equals_single_with_wide_kx0_plane(handle, prec):
    from openatom_amd.ortho import Batch
    shapes = [(96, 3000, 800), (64, 2048, 5), (40, 1200, 300)]
    psis = [make_states(n, p, eps=1e-3, seed=20 + i, nkx0=k, device=DEV, rng="torch")
            for i, (n, p, k) in enumerate(shapes)]
    b = Batch(handle)
    b.plan(psis, nkx0=[k for _, _, k in shapes], nspin=2, tol=1e-12, maxit=20, prec=prec)
    b.launch()
    torch.cuda.synchronize()
    for i, (n, p, k) in enumerate(shapes):
        S, T, out, _, _ = handle.ortho_step(psis[i], nkx0=k, nspin=2, tol=1e-12, maxit=20,
                                            prec=prec)
        torch.cuda.synchronize()
        assert torch.equal(S, b.S[i]), i
        assert torch.equal(out, b.outs[i]), i
    b.close()
