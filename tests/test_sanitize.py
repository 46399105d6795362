"""The host code under AddressSanitizer, UndefinedBehaviorSanitizer and
LeakSanitizer (CPU).

tests/cpp/sanitize_driver.cpp is compiled together with the oracle
(oracle/qhg_oracle.c), the species configuration (qhg4_amd/csrc/config.cpp)
and the QDF I/O (qhg4_amd/csrc/qdf_io.cpp) with
-fsanitize=address,undefined -fno-sanitize-recover=all.  It runs five action
sets on the oracle with 1 and 4 OpenMP threads (the tutorial model with
ATanDeath; the OoA model with genetics and RandomPair, RandomPair with
Navigate and a navigation event, AssortativePairHamming or
AssortativePair), every MD5 padding case, the
proteome reader at 12 genome sizes, 200 RankTable pairings, the
configuration's error cases, and a QDF write / read round trip with its
error cases.  Any sanitizer report ends the program with a failure.
"""
import os
import subprocess

import pytest

from tests.cpp.build_driver import HDF5, build_sanitize

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

pytestmark = pytest.mark.skipif(not os.path.exists(os.path.join(HDF5, "include", "hdf5.h")) or
                                not os.path.exists(os.path.join(ROOT, "qhg4_amd", "_hdf5")),
                                reason="needs the HDF5 headers and the package build's HDF5 link")


def test_host_code_under_sanitizers(tmp_path):
    exe = build_sanitize()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "all checks passed" in r.stdout
    for marker in ("ERROR: AddressSanitizer", "runtime error:", "ERROR: LeakSanitizer"):
        assert marker not in out, out[-4000:]
