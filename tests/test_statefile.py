"""State-file reader/writer (host calls of liboa_ortho.so, no GPU needed).

Pins: the reference's own k-point state files (tests/golden/gwbse_kpt0_state
{1,2}.out.gz, parsed independently by tests/golden/make_gwbse.py into
gwbse_kpt0.npz); for the gamma-point layout, a half-sphere file written here
from a conjugate-symmetric full sphere, whose expanded kx = 0 plane must be the
full plane in (ky, kz) order with c(-G) = conj c(G), and whose path-metric
overlap must equal the full-sphere complex overlap; byte-identical rewrites.
"""
import os
import struct

import numpy as np
import pytest

from openatom_amd import _lib
from openatom_amd.statefile import (BINARY, GZIP_BINARY, GZIP_TEXT, TEXT, State, read_state,
                                    write_state)

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module", autouse=True)
def _library():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("liboa_ortho.so not built")


def kpoint_path(j):
    return os.path.join(GOLDEN, f"gwbse_kpt0_state{j}.out")  # format 2 appends .gz


@pytest.mark.parametrize("j", [1, 2])
def test_reads_reference_kpoint_files(j):
    d = np.load(os.path.join(GOLDEN, "gwbse_kpt0.npz"))
    s = read_state(kpoint_path(j), GZIP_TEXT, double_pack=False)
    assert np.array_equal(s.coef, d["psi"][j - 1])
    assert np.array_equal(s.kvec, d["gvec"])
    assert s.grid == tuple(int(x) for x in d["grid"]) and s.nkx0 == 0


@pytest.mark.parametrize("fmt", [TEXT, BINARY, GZIP_TEXT, GZIP_BINARY])
def test_kpoint_round_trip(tmp_path, fmt):
    s = read_state(kpoint_path(1), GZIP_TEXT, double_pack=False)
    p = str(tmp_path / "state1.out")
    write_state(p, s, fmt, double_pack=False)
    r = read_state(p, fmt, double_pack=False)
    # the writer orders the points by (kx, ky, kz)
    order = np.lexsort((s.kvec[:, 2], s.kvec[:, 1], s.kvec[:, 0]))
    assert np.array_equal(r.kvec, s.kvec[order])
    want = s.coef[order]
    if fmt in (TEXT, GZIP_TEXT):  # "%g" keeps six significant digits, as the reference's
        want = np.array([complex(float(f"{z.real:g}"), float(f"{z.imag:g}")) for z in want])
    assert np.array_equal(r.coef, want)


# ---- gamma point: half-sphere files ---------------------------------------

def full_sphere(radius, seed):
    """All G with |G|^2 <= radius^2 and a conjugate-symmetric coefficient set
    (a real function in real space): dict G -> c, c(-G) = conj c(G)."""
    rng = np.random.default_rng(seed)
    r = int(radius)
    pts = [(x, y, z) for x in range(-r, r + 1) for y in range(-r, r + 1)
           for z in range(-r, r + 1) if x * x + y * y + z * z <= radius * radius]
    c = {}
    for g in sorted(pts):
        m = (-g[0], -g[1], -g[2])
        if m in c:
            c[g] = np.conj(c[m])
        elif g == (0, 0, 0):
            c[g] = complex(rng.standard_normal(), 0.0)
        else:
            c[g] = complex(rng.standard_normal(), rng.standard_normal())
    return c


def half_sphere_records(c):
    """The file order of a gamma-point state: the upper half of the kx = 0
    plane ((ky, kz) > (0, 0)), then kx > 0, each by (kx, ky, kz); G = 0 last."""
    upper = sorted(g for g in c if g[0] == 0 and (g[1], g[2]) > (0, 0))
    rest = sorted(g for g in c if g[0] > 0)
    return [(g, c[g]) for g in upper + rest] + [((0, 0, 0), c[(0, 0, 0)])]


def write_binary(path, recs, grid=(16, 16, 16)):
    with open(path, "wb") as f:
        f.write(struct.pack("4i", len(recs), *grid))
        for g, z in recs:
            f.write(struct.pack("2d3i", z.real, z.imag, *g))


def test_gamma_expansion_is_the_full_kx0_plane(tmp_path):
    c = full_sphere(4.3, seed=1)
    recs = half_sphere_records(c)
    p = str(tmp_path / "g.bin")
    write_binary(p, recs)
    s = read_state(p, BINARY, double_pack=True)
    nhalf = sum(1 for g, _ in recs if g[0] == 0) - 1
    assert len(s.coef) == len(recs) + nhalf and s.nkx0 == 2 * nhalf + 1
    plane = s.kvec[:s.nkx0]
    assert (plane[:, 0] == 0).all() and (s.kvec[s.nkx0:, 0] > 0).all()
    assert [tuple(g) for g in plane] == sorted(g for g in c if g[0] == 0)  # whole plane, in order
    for g, z in zip(s.kvec, s.coef):
        assert z == c[tuple(g)]  # the rebuilt half carries c(-G) = conj c(G)


def test_gamma_layout_gives_the_full_sphere_overlap(tmp_path, oracle):
    """The path metric on the expanded layout (weight 1 on the nkx0 leading
    points, 2 elsewhere: oao_overlap with alpha = 2 and the 1/sqrt2 kx = 0
    scaling) equals the complex overlap over the full sphere."""
    cs = [full_sphere(4.3, seed=s) for s in (2, 3, 4)]
    block = []
    for i, c in enumerate(cs):
        p = str(tmp_path / f"s{i}.bin")
        write_binary(p, half_sphere_records(c))
        st = read_state(p, BINARY, double_pack=True)
        block.append(st.coef)
    psi = np.ascontiguousarray(np.stack(block).view(np.float64))
    S = oracle.overlap(psi, st.nkx0)
    keys = sorted(cs[0])
    F = np.array([[c[g] for g in keys] for c in cs])
    S_full = (F.conj() @ F.T).real
    assert np.abs(S - S_full).max() < 1e-12 * np.abs(S_full).max()


@pytest.mark.parametrize("fmt", [BINARY, GZIP_BINARY])
def test_gamma_rewrite_is_byte_identical(tmp_path, fmt):
    recs = half_sphere_records(full_sphere(3.7, seed=5))
    src = str(tmp_path / "a.bin")
    write_binary(src, recs)
    s = read_state(src, BINARY, double_pack=True)
    out = str(tmp_path / "b.bin")
    write_state(out, State(s.coef, s.kvec, (16, 16, 16), s.nkx0), fmt, double_pack=True)
    if fmt == GZIP_BINARY:
        import gzip
        data = gzip.open(out + ".gz", "rb").read()
    else:
        data = open(out, "rb").read()
    assert data == open(src, "rb").read()


def test_gamma_text_matches_reference_format(tmp_path):
    recs = half_sphere_records(full_sphere(2.5, seed=6))
    src = str(tmp_path / "a.bin")
    write_binary(src, recs, grid=(8, 8, 8))
    s = read_state(src, BINARY, double_pack=True)
    out = str(tmp_path / "t.out")
    write_state(out, s, TEXT, double_pack=True)
    want = "%d %d %d %d\n" % (len(recs), 8, 8, 8) + "".join(
        "%g %g %d %d %d \n" % (z.real, z.imag, *g) for g, z in recs)
    assert open(out).read() == want


def test_errors(tmp_path):
    with pytest.raises(_lib.OaError):
        read_state(str(tmp_path / "missing"), TEXT)
    recs = half_sphere_records(full_sphere(2.5, seed=7))
    bad = str(tmp_path / "bad.bin")
    write_binary(bad, recs[1:])  # the upper half no longer starts at (0, 0, 1)
    with pytest.raises(_lib.OaError):
        read_state(bad, BINARY, double_pack=True)
    nog0 = str(tmp_path / "nog0.bin")
    write_binary(nog0, recs[:-1])  # no G = 0 record
    with pytest.raises(_lib.OaError):
        read_state(nog0, BINARY, double_pack=True)


# ---- device layout change and the step on states read from files ----------

def _gamma_block(tmp_path, nstates, radius, seed0):
    recs_all, mem = [], []
    for i in range(nstates):
        c = full_sphere(radius, seed=seed0 + i)
        norm = np.sqrt(sum(abs(z) ** 2 for z in c.values()))
        c = {g: z / norm for g, z in c.items()}  # full-sphere norm 1
        recs = half_sphere_records(c)
        p = str(tmp_path / f"s{i}.bin")
        write_binary(p, recs)
        recs_all.append(recs)
        mem.append(read_state(p, BINARY, double_pack=True))
    file_block = np.stack([np.array([z for _, z in r]) for r in recs_all]).view(np.float64)
    mem_block = np.stack([s.coef for s in mem]).view(np.float64)
    nhalf = sum(1 for g, _ in recs_all[0] if g[0] == 0) - 1
    return np.ascontiguousarray(file_block), np.ascontiguousarray(mem_block), nhalf, mem[0].nkx0


@pytest.mark.gpu
def test_device_kx0_expand_contract(tmp_path, handle):
    import torch
    from openatom_amd.statefile import kx0_contract, kx0_expand
    fb, mb, nhalf, _ = _gamma_block(tmp_path, 6, 5.2, seed0=20)
    f = torch.from_numpy(fb).cuda()
    m = kx0_expand(f, nhalf, handle=handle)
    back = kx0_contract(m, nhalf, handle=handle)
    torch.cuda.synchronize()
    assert np.array_equal(m.cpu().numpy(), mb)
    assert np.array_equal(back.cpu().numpy(), fb)


@pytest.mark.gpu
def test_step_on_states_read_from_files(tmp_path, handle):
    """Gamma-point states written as half-sphere files, read back into the
    memory layout, moved to the device and orthonormalised: T psi is
    orthonormal in the full-sphere metric that nkx0 encodes."""
    import torch
    _, mb, _, nkx0 = _gamma_block(tmp_path, 12, 5.2, seed0=40)
    psi = torch.from_numpy(mb).cuda()
    S, T, out, it, err = handle.ortho_step(psi, nkx0=nkx0, nspin=2, tol=1e-14, maxit=60)
    S2 = handle.overlap_sym(out, nkx0=nkx0, prec="fp64")
    torch.cuda.synchronize()
    assert (S2 - torch.eye(12, dtype=torch.float64, device="cuda")).abs().max().item() < 1e-12
