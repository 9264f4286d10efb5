"""Serving loop on the GPU: every request's result in the mix equals the
single-instance step on the same coefficients (tolerance: exact), and the
loop's accounting holds."""
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.serve import GpuExecutor, SessionSpec, serve  # noqa: E402
from openatom_amd.workload import make_states, md_perturbation, md_step_  # noqa: E402


def test_results_same_alone_and_in_mix(handle):
    steps = 4
    specs = [SessionSpec(0, 0.0, 64, 4096, steps, 11),
             SessionSpec(1, 0.0, 96, 6144, steps, 12),
             SessionSpec(2, 0.002, 128, 8192, steps, 13),
             SessionSpec(3, 0.004, 64, 4096, steps, 14),
             SessionSpec(4, 0.03, 160, 10240, steps, 15)]
    ex = GpuExecutor(handle, prec="int8x5", tol=1e-12, maxit=12, record=True)
    r = serve(specs, ex, budget_bytes=8 * 2**30)
    ex.release()
    assert r["sessions_finished"] == len(specs)
    assert r["requests_completed"] == steps * len(specs)
    assert r["mean_batch"] > 1.0  # sessions overlapped on the device
    for sp in specs:
        psi = make_states(sp.n, sp.npw, eps=1e-3, seed=sp.seed, device="cuda", rng="torch")
        P = md_perturbation(sp.n, sp.npw, seed=sp.seed * 1000 + 1)
        for k in range(steps):
            S, T, out, it, err = handle.ortho_step(psi, nkx0=1, nspin=2, tol=1e-12, maxit=12,
                                                   prec="int8x5")
            torch.cuda.synchronize()
            assert torch.equal(out, ex.records[sp.sid][k]), (sp.sid, k)
            psi = torch.empty_like(out)
            md_step_(psi, out, P, 1e-3, k + 1)


def test_serving_results_stay_orthonormal(handle):
    specs = [SessionSpec(i, 0.001 * i, n, 64 * n, 3, 40 + i)
             for i, n in enumerate((128, 256, 128, 512))]
    ex = GpuExecutor(handle, prec="int8x5", tol=1e-12, maxit=12, record=True)
    r = serve(specs, ex, budget_bytes=16 * 2**30)
    ex.release()
    assert r["sessions_finished"] == 4
    for sp in specs:
        out = ex.records[sp.sid][-1]
        S = handle.overlap_sym(out, nkx0=1, prec="fp64")
        eye = torch.eye(sp.n, dtype=torch.float64, device="cuda")
        assert (S - eye).abs().max().item() < 1e-10
