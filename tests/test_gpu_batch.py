"""Batched independent instances (grouped persistent launches, hipGraph
replay, tile autotuner) against the single-instance path.

The digit products are exact and the chunking depends on (n, npw) only, and
the FP64 S->T GEMMs keep the same k order for every tile, so each instance of
a batch must reproduce oa_ortho_step on that instance alone bit for bit, for
every tile choice and for graph replays."""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd.ortho import Batch  # noqa: E402
from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"


def uneven(count, seed, nmax=256, npw_max=8192):
    rng = np.random.default_rng(seed)
    ns = rng.choice(np.arange(64, nmax + 1, 32), size=count)
    npws = (rng.integers(1024, npw_max + 1, size=count) // 256) * 256
    eps = 10 ** rng.uniform(-4, -2, size=count)
    psis = [make_states(int(n), int(p), eps=float(e), seed=100 + i, device=DEV, rng="torch")
            for i, (n, p, e) in enumerate(zip(ns, npws, eps))]
    return psis


def alone(handle, psi, prec, nspin=2, tol=1e-12, maxit=20):
    S, T, out, it, err = handle.ortho_step(psi, nkx0=1, nspin=nspin, tol=tol, maxit=maxit,
                                           prec=prec)
    torch.cuda.synchronize()
    return S, T, out, it


@pytest.mark.parametrize("prec", ["int8x5", "fp16x5", "fp64"])
def test_batch_equals_alone(handle, prec):
    psis = uneven(7, seed=1)
    b = Batch(handle)
    b.plan(psis, nkx0=1, nspin=2, tol=1e-12, maxit=20, prec=prec)
    b.run()
    reps = b.reports()
    for i, psi in enumerate(psis):
        S, T, out, it = alone(handle, psi, prec)
        assert reps[i][0] == it and reps[i][1] == 1 and reps[i][2] == 0, (i, reps[i], it)
        assert torch.equal(b.S[i], S), (prec, i, psi.shape)
        assert torch.equal(b.T[i], T), (prec, i)
        assert torch.equal(b.outs[i], out), (prec, i)
    b.close()


@pytest.mark.parametrize("prec", ["int8x5", "fp64"])
def test_batch_graph_replay(handle, prec):
    psis = uneven(5, seed=2)
    b = Batch(handle)
    b.plan(psis, nkx0=1, nspin=2, tol=1e-12, maxit=12, prec=prec)
    b.run()
    torch.cuda.synchronize()
    ref = [o.clone() for o in b.outs]
    refT = [t.clone() for t in b.T]
    for rep in range(3):
        for o in b.outs:
            o.fill_(float("nan"))
        for t in b.T:
            t.zero_()
        b.launch()
        torch.cuda.synchronize()
        for i in range(len(psis)):
            assert torch.equal(b.outs[i], ref[i]), (rep, i)
            assert torch.equal(b.T[i], refT[i]), (rep, i)
        # eager work between replays
        _ = (psis[0] * 2).sum().item()
    b.close()


def test_batch_tile_choices_agree(handle):
    """Tile 128 and tile 64 digit GEMMs (and WM 2 / 4 S->T GEMMs) give the same
    bits; the tuner's table round-trips through text."""
    psis = uneven(6, seed=3)
    b = Batch(handle)
    buckets = sorted({(p.shape[0] + 31) // 32 for p in psis})
    outs = []
    for tile, wm in ((128, 4), (64, 2)):
        b.set_table("".join(f"{k} {tile} {wm}\n" for k in buckets))
        b.plan(psis, nkx0=1, nspin=2, tol=1e-12, maxit=12, prec="int8x5")
        assert all(b.config(i) == (tile, wm) for i in range(len(psis)))
        b.run()
        torch.cuda.synchronize()
        outs.append(([o.clone() for o in b.outs], [s.clone() for s in b.S]))
    for i in range(len(psis)):
        assert torch.equal(outs[0][0][i], outs[1][0][i]), i
        assert torch.equal(outs[0][1][i], outs[1][1][i]), i
    b.tune(reps=1)
    tab = b.table()
    assert len(tab.strip().splitlines()) == len(buckets)
    b.run()
    torch.cuda.synchronize()
    for i in range(len(psis)):
        assert torch.equal(b.outs[i], outs[0][0][i]), i
    b.close()


def test_batch_mixed_maxit_and_nonconvergence(handle):
    """An instance that hits its own invsqr_max_iter reports
    OA_ERR_NOT_CONVERGED without disturbing the others."""
    psis = uneven(3, seed=4)
    b = Batch(handle)
    b.plan(psis, nkx0=1, nspin=2, tol=[1e-12, 1e-30, 1e-12], maxit=[20, 2, 20], prec="int8x5")
    b.run()
    r = b.reports()
    assert r[1][0] == 2 and r[1][2] == -2
    assert r[0][2] == 0 and r[2][2] == 0
    S, T, out, it = alone(handle, psis[2], "int8x5")
    assert torch.equal(b.outs[2], out)
    b.close()


def test_handle_close_closes_its_batches():
    """A batch made on a handle is destroyed with it: closing the handle
    first (as a session teardown does) must leave nothing for the batch's
    own close or finalizer to touch."""
    from openatom_amd.ortho import Handle
    h = Handle(0)
    b = Batch(h)
    psi = make_states(32, 512, eps=1e-3, seed=1, device=DEV, rng="torch")
    b.plan([psi], nkx0=1, nspin=2, tol=1e-12, maxit=10, prec="int8x5")
    b.launch()
    torch.cuda.synchronize()
    h.close()
    assert b.b is None
    b.close()
    del b
    with pytest.raises(ValueError):
        Batch(h)
