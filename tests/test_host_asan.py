"""Host AddressSanitizer runs (CPU only) of the host C/C++ code that parses
or loops over caller buffers: the oracle's loop nests and plane splits
(oracle/oa_oracle.c) and the state-file reader/writer
(openatom_amd/csrc/oa_statefile.hip, host side; no GPU call is made).

Each driver under tests/asan/ is compiled into a temporary directory with the
sanitizer on the host code only (gcc -fsanitize=address,undefined for the C
oracle; hipcc -Xarch_host -fsanitize=address for the HIP source, so no GPU
code is built with it) and must exit 0 with no sanitizer report.
"""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN = os.path.join(ROOT, "tests", "asan")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd, env=None, timeout=300):
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, timeout=timeout)
    return r.returncode, r.stdout + r.stderr


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not installed")
def test_oracle_under_asan(tmp_path):
    exe = str(tmp_path / "oracle_asan")
    rc, log = _run(["gcc", "-O1", "-g", "-fsanitize=address,undefined",
                    "-fno-sanitize-recover=undefined", "-fno-omit-frame-pointer", "-fopenmp",
                    "-ffp-contract=off", os.path.join(ROOT, "oracle", "oa_oracle.c"),
                    os.path.join(ASAN, "oracle_driver.c"), "-lm", "-o", exe])
    assert rc == 0, log
    env = dict(os.environ, OMP_NUM_THREADS="2", ASAN_OPTIONS="detect_leaks=1:abort_on_error=0")
    rc, log = _run([exe], env=env)
    assert rc == 0 and "oracle_driver: ok" in log and "ERROR: AddressSanitizer" not in log, log


@pytest.mark.skipif(not os.path.exists(HIPCC), reason="hipcc not installed")
def test_statefile_parser_under_asan(tmp_path):
    exe = str(tmp_path / "statefile_asan")
    rc, log = _run([HIPCC, "-O1", "-g", "-std=c++17", "--offload-arch=gfx950",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer",
                    "-I", os.path.join(ROOT, "include"),
                    os.path.join(ROOT, "openatom_amd", "csrc", "oa_statefile.hip"),
                    os.path.join(ASAN, "statefile_driver.cpp"), "-lz", "-o", exe])
    assert rc == 0, log
    golden = os.path.join(ROOT, "tests", "golden", "gwbse_kpt0_state1.out")
    work = tmp_path / "files"
    work.mkdir()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    rc, log = _run([exe, golden, str(work)], env=env)
    assert rc == 0 and "statefile_driver: ok" in log and "AddressSanitizer" not in log, log
