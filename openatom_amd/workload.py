"""Synthetic plane-wave coefficient blocks for the orthonormalization path.

Generator (the workload every benchmark and parity test uses):

  * i.i.d. complex normal coefficients scaled by the envelope
    f(G) = 1 / (1 + |G|^2)^2 with |G_k| = (k / npw)^(1/3) * G_max, G_max = 10,
    coefficient k = 0 being G = 0 (forced real, as at the gamma point);
  * orthonormalised in FP64 under the path's metric
    <a, b> = 2 Re sum_{k >= nkx0} a_k^* b_k + Re sum_{k < nkx0} a_k^* b_k
    (the overlap the symmetric PairCalculator computes), by CholeskyQR2 --
    Gram-Schmidt's triangular factor obtained through two Cholesky passes;
  * perturbed by eps times a complex normal block with the same envelope,
    each row scaled to metric norm eps, so S = I + O(eps).

Rows are returned as the row-major [n][2*npw] interleaved (re, im) layout the
reference collates for DGEMM.  Random draws use numpy's PCG64 with the given
seed, so a seed names the same block on every machine; the orthonormalisation
runs on the requested torch device (plumbing only: no path arithmetic).
"""
from __future__ import annotations

import numpy as np
import torch

G_MAX = 10.0


def envelope(npw: int, gmax: float = G_MAX) -> np.ndarray:
    k = np.arange(npw, dtype=np.float64)
    G = (k / npw) ** (1.0 / 3.0) * gmax
    return 1.0 / (1.0 + G * G) ** 2


def metric_weights(npw: int, nkx0: int = 1) -> np.ndarray:
    """Per-double weights of the path metric over the interleaved layout."""
    w = np.full(2 * npw, 2.0)
    w[: 2 * nkx0] = 1.0
    return w


_DEV_CACHE: dict = {}


def _on_device(key, make, device) -> torch.Tensor:
    """Per-device cache of constant vectors (envelope, metric weights), so
    repeated draws do not copy them from host memory each time."""
    k = (key, str(torch.device(device)))
    t = _DEV_CACHE.get(k)
    if t is None:
        t = _DEV_CACHE[k] = torch.from_numpy(make()).to(device)
    return t


def _complex_normal_block(rng, n: int, npw: int, device) -> torch.Tensor:
    env = _on_device(("env", npw), lambda: np.repeat(envelope(npw), 2), device)
    if isinstance(rng, np.random.Generator):
        Z = torch.from_numpy(rng.standard_normal((n, 2 * npw))).to(device)
    else:
        Z = torch.randn((n, 2 * npw), generator=rng, dtype=torch.float64, device=device)
    Z *= env[None, :]
    Z[:, 1] = 0.0  # G = 0 coefficient is real
    return Z


def _cholqr2(Z: torch.Tensor, w: torch.Tensor) -> torch.Tensor:
    # cholesky_ex: no host-side check of the factorisation's info flag, so
    # generation on a side stream never stalls the host (S is a Gram matrix
    # of a full-rank random block, positive definite by construction)
    for _ in range(2):
        S = (Z * w) @ Z.T
        L, _info = torch.linalg.cholesky_ex(S)
        eye = torch.eye(L.shape[0], dtype=L.dtype, device=L.device)
        Linv = torch.linalg.solve_triangular(L, eye, upper=False)
        Z = Linv @ Z
    return Z


def make_states(n: int, npw: int, eps: float = 1e-3, seed: int = 0, nkx0: int = 1,
                device: str | torch.device = "cpu", rng: str = "numpy") -> torch.Tensor:
    """[n][2*npw] float64 block on `device`, S = I + O(eps) in the path metric.
    rng="numpy" draws with numpy's PCG64 (host, machine independent);
    rng="torch" draws with torch's Philox generator on `device` (fast, for
    the large benchmark blocks)."""
    if n > npw:
        raise ValueError("need n <= npw for an orthonormal block")
    if rng == "numpy":
        gen = np.random.default_rng(seed)
    else:
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
    Z = _complex_normal_block(gen, n, npw, device)
    w = torch.from_numpy(metric_weights(npw, nkx0)).to(device)
    Q = _cholqr2(Z, w)
    del Z
    if eps != 0.0:
        P = _complex_normal_block(gen, n, npw, device)
        nrm = torch.sqrt(((P * w) * P).sum(dim=1, keepdim=True))
        Q += P * (eps / nrm)
        del P
    return Q.contiguous()


def make_kpoint_states(n: int, npw: int, eps: float = 1e-3, seed: int = 0,
                       device: str | torch.device = "cpu", rng: str = "numpy") -> torch.Tensor:
    """[n][2*npw] float64 block of n complex states over the full G set of a
    k-point (no double packing): envelope-scaled complex normal coefficients,
    orthonormalised under <a, b> = sum_k conj(a_k) b_k by complex CholeskyQR2,
    plus eps times an envelope-shaped complex block with rows of norm eps, so
    that S = psi psi^H = I + O(eps).  rng="numpy" draws with numpy's PCG64,
    rng="torch" with torch's generator on `device` (the large blocks)."""
    if n > npw:
        raise ValueError("need n <= npw for an orthonormal block")
    if rng == "numpy":
        gen = np.random.default_rng(seed)
    else:
        gen = torch.Generator(device=device)
        gen.manual_seed(seed)
    env = _on_device(("env1", npw), lambda: envelope(npw), device)

    def block():
        if rng == "numpy":
            Z = torch.from_numpy(gen.standard_normal((n, 2 * npw))).to(device)
        else:
            Z = torch.randn((n, 2 * npw), generator=gen, dtype=torch.float64, device=device)
        return torch.view_as_complex(Z.reshape(n, npw, 2)) * env[None, :]

    Q = block()
    for _ in range(2):
        # invert the n x n factor and apply it with a GEMM: a triangular solve
        # with npw right-hand sides needs a library workspace that large
        # blocks exceed
        L, _info = torch.linalg.cholesky_ex(Q @ Q.conj().T)
        eye = torch.eye(n, dtype=L.dtype, device=L.device)
        Q = torch.linalg.solve_triangular(L, eye, upper=False) @ Q
    if eps != 0.0:
        P = block()
        Q = Q + P * (eps / torch.linalg.vector_norm(P, dim=1, keepdim=True))
    return torch.view_as_real(Q.contiguous()).reshape(n, 2 * npw).contiguous()


def md_perturbation(n: int, npw: int, seed: int, nkx0: int = 1,
                    device: str | torch.device = "cuda") -> torch.Tensor:
    """An envelope-shaped complex normal block with rows of unit metric norm,
    drawn once per trajectory; the serving loop's simulated MD step is
    psi_{k+1} = (T psi)_k + (-1)^k eps P (one fused pass over the block)."""
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    P = _complex_normal_block(gen, n, npw, device)
    w = _on_device(("w", npw, nkx0), lambda: metric_weights(npw, nkx0), device)
    P /= torch.sqrt(((P * w) * P).sum(dim=1, keepdim=True))
    return P


def md_step_(psi: torch.Tensor, out: torch.Tensor, P: torch.Tensor, eps: float,
             step: int) -> torch.Tensor:
    """psi <- out + (-1)^step eps P, in place."""
    return torch.add(out, P, alpha=(eps if step % 2 == 0 else -eps), out=psi)


def perturb_(psi: torch.Tensor, eps: float, seed: int, nkx0: int = 1) -> torch.Tensor:
    """One simulated MD step on a resident block, in place: psi += eps times
    an envelope-shaped complex normal block (each row of metric norm eps),
    drawn with torch's generator seeded by `seed` on psi's device."""
    n, K = psi.shape
    npw = K // 2
    gen = torch.Generator(device=psi.device)
    gen.manual_seed(seed)
    P = _complex_normal_block(gen, n, npw, psi.device)
    w = _on_device(("w", npw, nkx0), lambda: metric_weights(npw, nkx0), psi.device)
    nrm = torch.sqrt(((P * w) * P).sum(dim=1, keepdim=True))
    psi += P * (eps / nrm)
    return psi


def make_force(n: int, npw: int, std: float = 0.1, seed: int = 1,
               device: str | torch.device = "cpu", rng: str = "numpy") -> torch.Tensor:
    """Force block F for the Lambda product: i.i.d. complex normal, std 0.1."""
    if rng == "numpy":
        F = np.random.default_rng(seed).standard_normal((n, 2 * npw)) * (std / np.sqrt(2.0))
        return torch.from_numpy(F).to(device)
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    return torch.randn((n, 2 * npw), generator=gen, dtype=torch.float64,
                       device=device) * (std / np.sqrt(2.0))


def step_flops(n: int, npw: int, iters: int) -> float:
    """Algorithmic flops of one ortho step: S over the upper triangle
    (2n(n+1)npw), T*psi (4 n^2 npw) and S->T (6 n^3 per iteration)."""
    return 2.0 * n * (n + 1) * npw + 4.0 * n * n * npw + iters * 6.0 * n ** 3


def step_bytes(n: int, npw: int) -> float:
    """Minimum HBM bytes of one step: read psi for S, read psi and write T*psi."""
    return 48.0 * n * npw
