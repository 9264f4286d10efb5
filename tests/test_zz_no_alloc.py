"""oa_ortho_step_noalloc (the FP64 step whose buffers live in the library's
per-device scratch array) against oa_ortho_step, both called through the C
ABI: two calls in a row, the second on a smaller block, S, T and T psi within
1e-12 relative (the two differ only in how the overlap chunks its k sum), the
same iteration count, and the size limit and precision checks."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

from openatom_amd import _lib  # noqa: E402
from openatom_amd.workload import make_states  # noqa: E402

DEV = "cuda"


def rel(a, b):
    return float((a - b).abs().max() / b.abs().max())


@pytest.fixture(scope="module")
def capi():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    L = _lib.lib()
    h = ctypes.c_void_p()
    assert L.oa_create(ctypes.byref(h), 0) == 0
    yield L, h
    L.oa_destroy(h)


def step(L, h, fn, psi, nkx0=1, nspin=2, tol=1e-12, maxit=30, prec=0):
    n, K = psi.shape
    S = torch.full((n, n), float("nan"), dtype=torch.float64, device=DEV)
    T = torch.full_like(S, float("nan"))
    out = torch.full_like(psi, float("nan"))
    it, err = ctypes.c_int(-1), ctypes.c_double(-1.0)
    torch.cuda.synchronize()
    st = fn(h, n, K // 2, psi.data_ptr(), K, nkx0, nspin, tol, maxit, prec, S.data_ptr(),
            T.data_ptr(), out.data_ptr(), K, ctypes.byref(it), ctypes.byref(err))
    torch.cuda.synchronize()
    return st, S, T, out, it.value, err.value


def test_scratch_size(capi):
    L, _ = capi
    assert L.oa_noalloc_scratch_bytes() == 64 << 20


@pytest.mark.parametrize("shapes", [[(256, 16384), (96, 3000)], [(1024, 65536), (130, 2049)]])
def test_noalloc_matches_ortho_step(capi, shapes):
    L, h = capi
    for i, (n, npw) in enumerate(shapes):
        print(f"noalloc step n={n} npw={npw}", flush=True)
        psi = make_states(n, npw, eps=1e-3, seed=11 + i, device=DEV, rng="torch")
        want = step(L, h, L.oa_ortho_step, psi)
        got = step(L, h, L.oa_ortho_step_noalloc, psi)
        assert want[0] == 0 and got[0] == 0, (want[0], got[0])
        assert got[4] == want[4]
        for name, g, w in (("S", got[1], want[1]), ("T", got[2], want[2]),
                           ("T psi", got[3], want[3])):
            e = rel(g, w)
            assert e < 1e-12, (n, npw, name, e)
        del psi


def test_noalloc_async_matches(capi):
    """iters = err = NULL: the asynchronous form (all maxit guarded
    iterations on the stream, the report in the scratch array)."""
    L, h = capi
    psi = make_states(200, 6000, eps=1e-3, seed=5, device=DEV, rng="torch")
    n, K = psi.shape
    want = step(L, h, L.oa_ortho_step, psi, maxit=12)
    S = torch.empty((n, n), dtype=torch.float64, device=DEV)
    T = torch.empty_like(S)
    out = torch.empty_like(psi)
    torch.cuda.synchronize()
    st = L.oa_ortho_step_noalloc(h, n, K // 2, psi.data_ptr(), K, 1, 2, 1e-12, 12, 0,
                                 S.data_ptr(), T.data_ptr(), out.data_ptr(), K, None, None)
    assert st == 0
    assert L.oa_synchronize(h) == 0
    for g, w in ((S, want[1]), (T, want[2]), (out, want[3])):
        assert rel(g, w) < 1e-12


def test_noalloc_limits(capi):
    L, h = capi
    psi = make_states(64, 512, eps=1e-3, seed=1, device=DEV, rng="torch")
    st = step(L, h, L.oa_ortho_step_noalloc, psi, prec=_lib.prec_code("int8x5"))[0]
    assert st == _lib.OA_ERR_UNSUPPORTED
    # n = 1400: the S->T buffers alone (5 n^2 doubles) exceed the scratch array
    n = 1400
    big = torch.zeros((n, 4), dtype=torch.float64, device=DEV)
    S = torch.empty((n, n), dtype=torch.float64, device=DEV)
    out = torch.empty_like(big)
    st = L.oa_ortho_step_noalloc(h, n, 2, big.data_ptr(), 4, 1, 2, 1e-12, 5, 0, S.data_ptr(),
                                 S.data_ptr(), out.data_ptr(), 4, None, None)
    assert st == _lib.OA_ERR_CAPACITY
    assert L.oa_ortho_step_noalloc(h, 8, 0, big.data_ptr(), 4, 0, 2, 1e-12, 5, 0,
                                   S.data_ptr(), S.data_ptr(), out.data_ptr(), 4, None,
                                   None) == 3
