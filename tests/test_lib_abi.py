"""The C-ABI library (CPU checks: no compute calls without a GPU)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd import _lib
from oracle.oracle import lib as olib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    return set(re.findall(r"\b(qhg_[a-z0-9_]+)\s*\(", txt))


def test_every_header_is_covered():
    assert sorted(fn for fn in os.listdir(os.path.join(ROOT, "include")) if fn.endswith(".h")) == \
        ["qhg_gpu.h", "qhg_qdf.h"]


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    decl = declared_symbols("qhg_gpu.h")
    assert decl, "no declarations found"
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert set(_lib.EXPORTS) == decl


def test_qdf_library_exports_every_declared_symbol():
    from qhg4_amd import qdf
    L = qdf.lib()
    decl = declared_symbols("qhg_qdf.h")
    for name in sorted(decl):
        assert hasattr(L, name), name
    assert set(qdf.EXPORTS) == decl


def test_abi_version_and_words():
    L = _lib.lib()
    assert L.qhg_gpu_abi_version() == 8
    assert L.qhg_gpu_genome_words(256) == 8 == abi.genome_words(256)
    assert L.qhg_gpu_genome_words(1) == 1


def test_struct_sizes_match_header():
    # field order / padding of the ctypes mirrors (qhg_params is the largest)
    assert C.sizeof(abi.Params) == 8 + 4 + 64 + 64 + 16 * 48 + 9 * 8 + 4 + 4 + 16 + 16 + 32 + 16 + \
        2 * 4 * 64 * 8 + 9 * 8 + 4 + 4 + 8 + 4 + 5 * 8 + 8 + 4 + 4 + 5 * 8  # 4 bytes of padding before the first double
    # qhg_nav_desc: int32 (+4 padding), four pointers, a double, int32 (+4), a pointer
    assert C.sizeof(abi.NavDesc) == 8 + 4 * 8 + 8 + 8 + 8
    assert C.sizeof(abi.Stats) == 17 * 8
    assert C.sizeof(abi.Tuning) == 16


def test_tuning_entry_points_validate_without_a_gpu():
    L = _lib.lib()
    t = abi.Tuning(64.0, 16, 0)
    assert L.qhg_gpu_set_tuning(None, C.byref(t)) == -1
    assert L.qhg_batch_set_tuning(None, C.byref(t)) == -1
    assert L.qhg_batch_autotune(None, 4, None, None) == -1
    assert L.qhg_gpu_device_bytes(None, None) == -1


@pytest.mark.parametrize("p,n", [(1e-3, 512), (1e-2, 512), (0.2, 64)])
def test_binomial_table_equals_oracle(p, n):
    """The product's host table (shared by every device launch) is bit-equal
    to the oracle's restatement."""
    a = (C.c_double * 64)()
    na = C.c_int32()
    assert _lib.lib().qhg_gpu_binomial_table(p, n, 1e-6, a, C.byref(na)) == 0
    b = (C.c_double * 64)()
    nb = olib().orc_bino_table(p, n, 1e-6, b, 64)
    assert na.value == nb
    assert list(a)[:nb] == list(b)[:nb]


def test_binomial_table_rejects_bad_args():
    a = (C.c_double * 64)()
    na = C.c_int32()
    assert _lib.lib().qhg_gpu_binomial_table(1.5, 10, 1e-6, a, C.byref(na)) != 0
    assert b"binomial" in _lib.lib().qhg_gpu_last_error()


def test_create_validates_arguments_before_touching_the_gpu():
    sc = scenarios.c1(width=8, height=8, n_agents=10)
    g = scenarios.grid_desc(sc["grid"])
    e = scenarios.env_desc({})
    h = C.c_void_p()
    bad = scenarios.grid_desc(sc["grid"])
    bad.num_cells = 0
    assert _lib.lib().qhg_gpu_create(C.byref(bad), C.byref(e), C.byref(sc["params"]), 0, C.byref(h)) == -1
    nbr = sc["grid"]["neighbors"].copy()
    nbr[0, 0] = 9999
    g2 = scenarios.grid_desc(dict(sc["grid"], neighbors=nbr))
    assert _lib.lib().qhg_gpu_create(C.byref(g2), C.byref(e), C.byref(sc["params"]), 0, C.byref(h)) == -1
    assert b"neighbour" in _lib.lib().qhg_gpu_last_error()
    del g


@pytest.mark.skipif(os.path.exists("/dev/kfd") and os.access("/dev/kfd", os.R_OK | os.W_OK),
                    reason="a GPU is present")
def test_no_gpu_fails_loudly():
    from qhg4_amd.population import GpuPopulation
    from qhg4_amd._lib import QhgError
    sc = scenarios.c1(width=8, height=8, n_agents=10)
    with pytest.raises(QhgError):
        GpuPopulation.from_scenario(sc)
