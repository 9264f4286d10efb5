"""Write the complex k-point fixtures from the reference's own state files.

Input: the 8 OpenAtom state files of k-points 0 and 1 that the reference's
GW/BSE test input holds (src_charm_driver/gw_bse/input/STATES_IN/
Spin.0_Kpt.{0,1}_Bead.0_Temper.0/state{1..8}.out.gz; format: a header line
`ncoef nkx nky nkz`, then one `re im kx ky kz` line per plane wave, the full
sphere with G = 0 first).  Expected output: the Hermitian overlap
S = Psi^H Psi by the reference's ZGEMM compiled from src_mathlib/zgemm.f
(oracle/_ref/librefblas.so): ZGEMM('C', 'N', 8, 8, ncoef, 1, A, ncoef, A,
ncoef, 0, S, 8) with A(:, j) = state j.  Also copies state1 and state2 of
k-point 0 unchanged (gwbse_kpt0_state{1,2}.out.gz) for the state-file
reader's tests.  Run from the repository root:
python tests/golden/make_gwbse.py
"""
from __future__ import annotations

import gzip
import os
import shutil
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as o  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src_charm_driver/gw_bse/input/STATES_IN"


def read_state(path: str):
    with gzip.open(path, "rt") as f:
        head = f.readline().split()
        ncoef = int(head[0])
        rows = [f.readline().split() for _ in range(ncoef)]
    c = np.array([float(r[0]) + 1j * float(r[1]) for r in rows], dtype=np.complex128)
    g = np.array([[int(r[2]), int(r[3]), int(r[4])] for r in rows], dtype=np.int32)
    return c, g, np.array([int(x) for x in head[1:4]], dtype=np.int32)


def main() -> None:
    o.build()
    assert o.ref_available(), "reference BLAS not built (needs the reference tree)"
    for kpt in (0, 1):
        d = os.path.join(SRC, f"Spin.0_Kpt.{kpt}_Bead.0_Temper.0")
        states, gvec, grid = [], None, None
        for j in range(1, 9):
            c, g, grid = read_state(os.path.join(d, f"state{j}.out.gz"))
            if gvec is None:
                gvec = g
            assert np.array_equal(g, gvec), "states of one k-point share their G list"
            states.append(c)
        psi = np.ascontiguousarray(np.stack(states))  # row j = state j = column j of A
        nst, ncoef = psi.shape
        S = np.zeros((nst, nst), dtype=np.complex128)  # column-major buffer: S_cm[j, i] = S(i, j)
        o.zgemm_cm("C", "N", nst, nst, ncoef, 1.0, psi, ncoef, psi, ncoef, 0.0, S, nst,
                   which="ref")
        np.savez_compressed(os.path.join(OUT, f"gwbse_kpt{kpt}.npz"), psi=psi, gvec=gvec,
                            grid=grid, S_cm=S)
        print(f"kpt {kpt}: {nst} states x {ncoef} plane waves, "
              f"max|S - I| = {np.abs(S.T - np.eye(nst)).max():.2e}")
    # two of the state files themselves, unchanged, for the state-file reader
    d = os.path.join(SRC, "Spin.0_Kpt.0_Bead.0_Temper.0")
    for j in (1, 2):
        shutil.copyfile(os.path.join(d, f"state{j}.out.gz"),
                        os.path.join(OUT, f"gwbse_kpt0_state{j}.out.gz"))


if __name__ == "__main__":
    main()
