"""State files and the gamma-point kx = 0 plane layout, through the library's
C calls (oa_state_read / oa_state_write run on the host; oa_kx0_expand /
oa_kx0_contract on the device).

A state file (utility/util.C:1107-1330 readState, :1955-2180 writeStateFile)
is a header `ncoef nx ny nz` and one `re im kx ky kz` record per plane wave;
format 0 text, 1 binary, 2 gzip text, 3 gzip binary (2 and 3 read
`<path>.gz`).  With double packing (gamma point) the file holds the upper half
of the kx = 0 plane, the kx > 0 points and G = 0 last; in memory the whole
kx = 0 plane leads (para_grp_parse.C:178-264 flip_data_set), which is the
layout the path's nkx0 weighting expects.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import check

TEXT, BINARY, GZIP_TEXT, GZIP_BINARY = 0, 1, 2, 3


@dataclass
class State:
    coef: np.ndarray      # complex128 [npts]
    kvec: np.ndarray      # int32 [npts, 3]
    grid: tuple           # (nx, ny, nz)
    nkx0: int             # leading points with kx = 0 (double packing), else 0


def read_state(path: str, format: int = GZIP_TEXT, double_pack: bool = True) -> State:
    L = _lib.lib()
    npts, nkx0 = ctypes.c_int(0), ctypes.c_int(0)
    grid = (ctypes.c_int * 3)()
    st = L.oa_state_read(path.encode(), format, 1 if double_pack else 0, 0, None, None, None,
                         None, ctypes.byref(npts), ctypes.byref(nkx0), grid)
    if st not in (_lib.OA_SUCCESS, _lib.OA_ERR_CAPACITY):
        check(st, f"oa_state_read({path})")
    n = npts.value
    coef = np.empty(n, dtype=np.complex128)
    k = np.empty((3, n), dtype=np.int32)
    ip = ctypes.POINTER(ctypes.c_int)
    check(L.oa_state_read(path.encode(), format, 1 if double_pack else 0, n,
                          coef.ctypes.data_as(ctypes.c_void_p), k[0].ctypes.data_as(ip),
                          k[1].ctypes.data_as(ip), k[2].ctypes.data_as(ip),
                          ctypes.byref(npts), ctypes.byref(nkx0), grid),
          f"oa_state_read({path})")
    return State(coef, np.ascontiguousarray(k.T), tuple(grid), nkx0.value)


def write_state(path: str, state: State, format: int = TEXT, double_pack: bool = True) -> None:
    coef = np.ascontiguousarray(state.coef, dtype=np.complex128)
    k = np.ascontiguousarray(np.asarray(state.kvec, dtype=np.int32).T)
    grid = (ctypes.c_int * 3)(*state.grid)
    ip = ctypes.POINTER(ctypes.c_int)
    check(_lib.lib().oa_state_write(path.encode(), format, 1 if double_pack else 0, len(coef),
                                    coef.ctypes.data_as(ctypes.c_void_p),
                                    k[0].ctypes.data_as(ip), k[1].ctypes.data_as(ip),
                                    k[2].ctypes.data_as(ip), grid),
          f"oa_state_write({path})")


def read_states(paths, format: int = GZIP_TEXT, double_pack: bool = True,
                device: str | torch.device = "cuda"):
    """n state files of one k-point set -> the [n][2*npts] float64 block the
    path takes (rows = states), the shared k-vectors and nkx0."""
    states = [read_state(p, format, double_pack) for p in paths]
    kvec = states[0].kvec
    for s in states[1:]:
        if not np.array_equal(s.kvec, kvec):
            raise ValueError("the states of one block must share their k-vector list")
    block = np.stack([s.coef for s in states]).view(np.float64)
    return torch.from_numpy(np.ascontiguousarray(block)).to(device), kvec, states[0].nkx0


def kx0_expand(block: torch.Tensor, nhalf: int, handle=None) -> torch.Tensor:
    """File order (upper half of the kx = 0 plane: nhalf points, kx > 0, G = 0
    last) -> memory order [lower half | G = 0 | upper half | kx > 0], for every
    row of a device block."""
    from .ortho import _ld, _need, _ptr, default_handle
    h = handle or default_handle()
    _need(block, "block")
    n, nfile = block.shape[0], block.shape[1] // 2
    out = torch.empty((n, 2 * (nfile + nhalf)), dtype=torch.float64, device=block.device)
    check(_lib.lib().oa_kx0_expand(h.h, n, nfile, nhalf, _ptr(block), _ld(block), _ptr(out),
                                   _ld(out)), "oa_kx0_expand")
    return out


def kx0_contract(block: torch.Tensor, nhalf: int, handle=None) -> torch.Tensor:
    """Memory order -> file order (drops the rebuilt lower half)."""
    from .ortho import _ld, _need, _ptr, default_handle
    h = handle or default_handle()
    _need(block, "block")
    n, npts = block.shape[0], block.shape[1] // 2
    nfile = npts - nhalf
    out = torch.empty((n, 2 * nfile), dtype=torch.float64, device=block.device)
    check(_lib.lib().oa_kx0_contract(h.h, n, nfile, nhalf, _ptr(block), _ld(block), _ptr(out),
                                     _ld(out)), "oa_kx0_contract")
    return out
