"""The C-ABI library loads and exports every symbol include/oa_ortho.h
declares; host-side argument handling that needs no device."""
import ctypes
import subprocess

import numpy as np
import pytest

from openatom_amd import _lib


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    declared = _lib.declared_symbols()
    assert "dgemm_" in declared and "oa_ortho_step" in declared
    for name in declared:
        assert hasattr(L, name), name
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    exported = {l.split()[-1] for l in out.splitlines() if " T " in l}
    assert set(declared) <= exported


def test_library_is_gfx950_only():
    out = subprocess.run(["/opt/rocm/bin/roc-obj-ls", _lib.LIB_PATH], capture_output=True,
                         text=True)
    if out.returncode != 0:
        pytest.skip("roc-obj-ls unavailable")
    arches = {w for w in out.stdout.split() if "gfx" in w}
    assert arches and all("gfx950" in a for a in arches), arches


def test_prec_codes():
    assert _lib.prec_code("fp64") == 0
    assert _lib.prec_code("int8x5") == (1 | (5 << 8))
    assert _lib.prec_code("fp16") == (2 | (3 << 8))
    assert _lib.prec_code("bf16x4") == (3 | (4 << 8))
    with pytest.raises(ValueError):
        _lib.prec_code("fp8")


def test_status_strings():
    assert "iteration limit" in _lib.status_string(_lib.OA_ERR_NOT_CONVERGED)
    assert "invalid argument" in _lib.status_string(3)


def test_dgemm_invalid_arguments_return_silently():
    """dgemm.f's XERBLA is an empty body: bad arguments leave C untouched;
    this path returns before any device work."""
    L = _lib.lib()
    f = L.dgemm_
    f.restype = None
    ci, cd = ctypes.c_int, ctypes.c_double
    dp = ctypes.POINTER(ctypes.c_double)
    A = np.ones(16)
    C = np.full(16, 5.0)
    for ta, tb, m, n, k, lda, ldb, ldc in [(b"X", b"N", 2, 2, 2, 2, 2, 2),
                                           (b"N", b"N", 4, 2, 2, 3, 2, 4),
                                           (b"N", b"N", 2, 2, 2, 2, 2, 1)]:
        f(ctypes.c_char_p(ta), ctypes.c_char_p(tb), ctypes.byref(ci(m)), ctypes.byref(ci(n)),
          ctypes.byref(ci(k)), ctypes.byref(cd(1.0)), A.ctypes.data_as(dp), ctypes.byref(ci(lda)),
          A.ctypes.data_as(dp), ctypes.byref(ci(ldb)), ctypes.byref(cd(0.0)),
          C.ctypes.data_as(dp), ctypes.byref(ci(ldc)))
        assert np.all(C == 5.0)
    # m == 0: quick return
    f(ctypes.c_char_p(b"N"), ctypes.c_char_p(b"N"), ctypes.byref(ci(0)), ctypes.byref(ci(2)),
      ctypes.byref(ci(2)), ctypes.byref(cd(1.0)), A.ctypes.data_as(dp), ctypes.byref(ci(1)),
      A.ctypes.data_as(dp), ctypes.byref(ci(2)), ctypes.byref(cd(0.0)), C.ctypes.data_as(dp),
      ctypes.byref(ci(1)))
    assert np.all(C == 5.0)


def test_handle_api_argument_checks_without_device():
    """oa_* entry points validate a null handle before touching a device."""
    L = _lib.lib()
    assert L.oa_overlap_sym(None, 4, 4, None, 8, 1, 2.0, 0, None, 0.0, None) == _lib.OA_ERR_NO_DEVICE
    assert L.oa_destroy(None) == 0
