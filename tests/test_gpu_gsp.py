"""GPU parity of the GSpace side of the Lambda round trip (oa_lambda_force,
oa_scale_kx0 and the GSpacePlane host mirror) against the CPU oracle.

Tolerances: FP64 1e-12 relative to the largest force entry (the backward
product accumulates in another order than the reference's DGEMM); int8x5
1e-6: Lambda is a general matrix, not I + O(eps), so its digit products get
no diagonal gain and carry the 4-slice truncation (about 1e-8 .. 1e-7 of the
largest entry).  Column weights are exact scalings, so the weighted and unweighted
columns are compared bit for bit with plain back-multiplies.
"""
import math

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.ortho import GSpacePlane, Ortho, PairCalculator  # noqa: E402
from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"


def rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / max(np.abs(b).max(), 1e-300))


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(DEV)


def states(n, npw, nkx0, double_pack=True, lsda=False, seed=0):
    psi = make_states(n, npw, eps=0.0, seed=seed, nkx0=nkx0 if double_pack else npw).numpy()
    return psi * (1.0 if lsda else math.sqrt(2.0))


@pytest.mark.parametrize("prec", ["fp64", "int8x5"])
@pytest.mark.parametrize("transM", [False, True])
def test_lambda_force_columns_exact(handle, prec, transM):
    """wkx0 = 1 is the plain back-multiply; with beta = 0 the kx = 0 columns
    equal a back-multiply with alpha * wkx0 and the rest one with alpha,
    bit for bit."""
    n, npw, nkx0 = 160, 3000, 37
    rng = np.random.default_rng(1)
    psi = dev(states(n, npw, nkx0, seed=2))
    M = dev(rng.standard_normal((n, n)) * 0.1)
    base = dev(rng.standard_normal((n, 2 * npw)))
    f1 = base.clone()
    handle.lambda_force(M, psi, f1, transM=transM, nkx0=nkx0, wkx0=1.0, alpha=-1.0, beta=1.0,
                        prec=prec)
    f2 = handle.backmul(M, psi, transM=transM, alpha=-1.0, beta=1.0, out=base.clone(), prec=prec)
    torch.cuda.synchronize()
    assert torch.equal(f1, f2)
    w = handle.lambda_force(M, psi, torch.empty_like(base), transM=transM, nkx0=nkx0, wkx0=0.5,
                            alpha=-1.0, beta=0.0, prec=prec)
    a = handle.backmul(M, psi, transM=transM, alpha=-1.0, prec=prec)
    b = handle.backmul(M, psi, transM=transM, alpha=-0.5, prec=prec)
    torch.cuda.synchronize()
    assert torch.equal(w[:, 2 * nkx0:], a[:, 2 * nkx0:])
    assert torch.equal(w[:, :2 * nkx0], b[:, :2 * nkx0])


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-12), ("int8x5", 1e-6)])
@pytest.mark.parametrize("mode,lsda,double_pack", [
    (0, False, True), (0, True, True), (1, False, True), (1, True, True), (0, False, False),
    (1, True, False)])
def test_unpack_lambda_vs_oracle(handle, oracle, prec, tol, mode, lsda, double_pack):
    n, npw, nkx0 = 192, 4096, 45
    psi_h = states(n, npw, nkx0, double_pack, lsda, seed=3)
    F_h = np.random.default_rng(4).standard_normal((n, 2 * npw))
    L_h = oracle.lambda_fwd(psi_h, F_h)
    if mode == 0 and lsda:
        L_h = 2.0 * L_h
    want = F_h.copy()
    oracle.gsp_unpack(oracle.backmul(L_h, psi_h, transT=True), want, nkx0=nkx0,
                      double_pack=double_pack, mode=mode, lsda=lsda)
    g = GSpacePlane(nkx0=nkx0, double_pack=double_pack, cp_min_opt=1, nspin=2 if lsda else 1,
                    prec=prec, handle=handle)
    got = g.unpackLambda(dev(L_h), dev(psi_h), dev(F_h), reduction=(mode == 1))
    torch.cuda.synchronize()
    assert rel(got.cpu().numpy(), want) < tol


@pytest.mark.parametrize("prec", ["fp64", "int8x5"])
@pytest.mark.parametrize("lsda", [False, True])
def test_minimisation_round_trip_projects_out_lambda(handle, prec, lsda):
    """Forward Lambda, Ortho's lsda doubling, backward with the GSpace
    weights, all on the device: the updated force has zero Lambda with the
    (orthonormal) states."""
    n, npw, nkx0 = 256, 8192, 60
    psi = dev(states(n, npw, nkx0, lsda=lsda, seed=5))
    F = dev(np.random.default_rng(6).standard_normal((n, 2 * npw)))
    pc = PairCalculator(symmetric=False, nkx0=nkx0, prec=prec, handle=handle)
    ortho = Ortho(nspin=2 if lsda else 1, handle=handle)
    g = GSpacePlane(nkx0=nkx0, cp_min_opt=1, nspin=2 if lsda else 1, prec=prec, handle=handle)
    L = ortho.acceptSectionLambda(pc.multiplyForward(psi, F))
    g.unpackLambda(L, psi, F)
    L1 = PairCalculator(symmetric=False, prec="fp64", handle=handle).multiplyForward(psi, F)
    torch.cuda.synchronize()
    bound = 1e-12 if prec == "fp64" else 1e-6
    assert L1.abs().max().item() < bound * L.abs().max().item()


@pytest.mark.parametrize("lsda", [False, True])
def test_dynamics_round_trip_vs_oracle(handle, oracle, lsda):
    """scaleForLambda -> Lambda -> S->T -> Gamma = T Lambda -> unpackGamma on
    the device against the oracle's GSpace/PairCalculator sequence; the
    resulting force is Lambda-free for orthonormal states."""
    n, npw, nkx0 = 128, 2048, 30
    psi_h = states(n, npw, nkx0, lsda=lsda, seed=7)
    F_h = np.random.default_rng(8).standard_normal((n, 2 * npw))
    nspin = 2 if lsda else 1
    g = GSpacePlane(nkx0=nkx0, cp_min_opt=0, nspin=nspin, handle=handle)
    psi, F = dev(psi_h), dev(F_h)
    g.scaleForLambda(psi, F)
    ps_h, Fs_h = psi_h.copy(), F_h.copy()
    ps_h[:, :2 * nkx0] *= 1.0 / math.sqrt(2.0)
    Fs_h[:, :2 * nkx0] *= math.sqrt(2.0)
    torch.cuda.synchronize()
    assert np.array_equal(psi.cpu().numpy(), ps_h) and np.array_equal(F.cpu().numpy(), Fs_h)
    sym = PairCalculator(symmetric=True, nkx0=nkx0, handle=handle)
    S = sym.multiplyForward(dev(psi_h))
    ortho = Ortho(nspin=nspin, handle=handle)
    T = ortho.start_calc(S)
    L = PairCalculator(symmetric=False, handle=handle).multiplyForward(psi, F)
    G = ortho.gamma(T, L)
    force = torch.empty_like(F)
    g.unpackGamma(G, T, psi, F, force)
    # oracle sequence on the same T, Lambda and Gamma
    T_h, G_h = T.cpu().numpy(), G.cpu().numpy()
    data = oracle.backmul(T_h, Fs_h, transT=False, alpha=-1.0)
    oracle.backmul(G_h, ps_h, transT=True, alpha=1.0, beta=1.0, out=data)
    want = np.zeros_like(F_h)
    oracle.gsp_unpack(data, want, nkx0=nkx0, double_pack=True, mode=2, lsda=lsda)
    got = force.cpu().numpy()
    assert rel(got, want) < 1e-12
    Lf = oracle.lambda_fwd(psi_h, got)
    assert np.abs(Lf).max() < 1e-11 * np.abs(L.cpu().numpy()).max()
    with pytest.raises(ValueError):
        g.unpackGamma(G, T, psi, F, F)


def test_scale_kx0_exact(handle):
    rng = np.random.default_rng(11)
    X_h = rng.standard_normal((77, 2 * 513))
    X = dev(X_h)
    handle.scale_kx0(X, 19, 1.0 / math.sqrt(2.0))
    want = X_h.copy()
    want[:, :38] *= 1.0 / math.sqrt(2.0)
    torch.cuda.synchronize()
    assert np.array_equal(X.cpu().numpy(), want)
    handle.scale_kx0(X, 513, 2.0)  # every point: the single-pack form
    want *= 2.0
    torch.cuda.synchronize()
    assert np.array_equal(X.cpu().numpy(), want)


def test_lambda_force_argument_errors(handle):
    from openatom_amd import _lib
    psi = torch.zeros((4, 20), dtype=torch.float64, device=DEV)
    M = torch.zeros((4, 4), dtype=torch.float64, device=DEV)
    L = _lib.lib()
    p = psi.data_ptr()
    assert L.oa_lambda_force(handle.h, 4, 10, M.data_ptr(), 0, p, 20, 11, 0.5, -1.0, 1.0, p,
                             20, 0) == 8
    assert L.oa_lambda_force(handle.h, 4, 10, M.data_ptr(), 0, p, 19, 1, 0.5, -1.0, 1.0, p,
                             20, 0) == 7
    assert L.oa_lambda_force(handle.h, 4, 10, M.data_ptr(), 0, p, 20, 1, 0.5, -1.0, 1.0, p,
                             18, 0) == 13
    assert L.oa_scale_kx0(handle.h, 4, 10, p, 20, 11, 2.0) == 6
