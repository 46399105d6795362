"""Pins of the oracle against the reference's own utility code.

oracle/_ref/libqhg_ref.so is compiled from the reference sources
(QHG4/utils/WELL512.cpp, QHG4/genes/GeneUtils.cpp, QHG4/utils/BinomialDist.cpp,
QHG4/utils/bino_tools.cpp and the string helpers) by `make -C oracle ref`.
The reference functions draw from a WELL512 generator; the wrappers report the
raw 32-bit words they consumed, and the oracle gets the same words, so every
comparison here is bit for bit.  Skipped where the reference tree is absent
(the library is built only where the reference sources are).
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import lib as olib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libqhg_ref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF),
                                reason="reference utilities not built (make -C oracle ref needs the reference tree)")

U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
F64P = C.POINTER(C.c_double)
I32P = C.POINTER(C.c_int32)


def ref():
    L = C.CDLL(REF)
    L.ref_well_draws.argtypes = [U32P, C.c_uint32, C.c_int, U32P, F64P, F64P, U32P, C.c_double, C.c_double,
                                 C.c_uint32, C.c_uint32, C.c_uint32]
    L.ref_multimask.argtypes = [U32P, C.c_int]
    L.ref_multimask.restype = C.c_uint64
    L.ref_crossover.argtypes = [U64P, U64P, C.c_int, C.c_int, U32P, U32P]
    L.ref_free_reco.argtypes = [U64P, U64P, C.c_int, U32P, U32P]
    L.ref_mutate_nucs.argtypes = [U64P, C.c_int, C.c_int, U32P, U32P]
    L.ref_bino_getn.argtypes = [C.c_double, C.c_int, C.c_double, F64P, C.c_int, I32P]
    L.ref_bino_getn.restype = C.c_int
    L.ref_gammaln.argtypes = [C.c_double]
    L.ref_gammaln.restype = C.c_double
    L.ref_beta.argtypes = [C.c_double, C.c_double, C.c_double]
    L.ref_beta.restype = C.c_double
    return L


def orc():
    L = olib()
    L.orc_wrandd.argtypes = [C.c_uint32]
    L.orc_wrandd.restype = C.c_double
    L.orc_wrandr.argtypes = [C.c_double, C.c_double, C.c_uint32]
    L.orc_wrandr.restype = C.c_double
    L.orc_wrandi.argtypes = [C.c_uint32, C.c_uint32, C.c_uint32, C.c_uint32]
    L.orc_wrandi.restype = C.c_uint32
    L.orc_multimask.argtypes = [U32P, C.c_int32]
    L.orc_multimask.restype = C.c_uint64
    L.orc_crossover.argtypes = [U64P, U64P, C.c_int32, U32P, C.c_int32]
    L.orc_free_reco.argtypes = [U64P, U64P, C.c_int32, U32P]
    L.orc_mutate_nucs.argtypes = [U64P, C.c_int32, C.c_int32, U32P, U32P]
    L.orc_bino_table.argtypes = [C.c_double, C.c_int32, C.c_double, F64P, C.c_int32]
    L.orc_bino_table.restype = C.c_int32
    L.orc_bino_getn.argtypes = [F64P, C.c_int32, C.c_double]
    L.orc_bino_getn.restype = C.c_int32
    L.orc_betai.argtypes = [C.c_double, C.c_double, C.c_double]
    L.orc_betai.restype = C.c_double
    return L


def p(a, t):
    return a.ctypes.data_as(t)


def states(seed, n):
    rng = np.random.default_rng(seed)
    return [rng.integers(0, 2**32, 16, dtype=np.uint64).astype(np.uint32) for _ in range(n)]


@pytest.mark.parametrize("a,b,ia,ib,s", [(0.0, 1.0, 0, 512, 2), (0.0, 17.0, 1, 4, 1), (-3.5, 2.25, 0, 1024, 2),
                                          (0.0, 63.0, 0, 65, 1), (10.0, 10.0, 0, 2**31, 3)])
def test_well512_conversions(a, b, ia, ib, s):
    """wrandd / wrandr / wrandi of the reference generator against the oracle's
    formulas applied to the same raw words (WELL512.h:33-39)."""
    R, O = ref(), orc()
    n = 4000
    for st in states(1, 4):
        raw = np.zeros(n, np.uint32)
        d = np.zeros(n)
        r = np.zeros(n)
        iv = np.zeros(n, np.uint32)
        R.ref_well_draws(p(st, U32P), 0, n, p(raw, U32P), p(d, F64P), p(r, F64P), p(iv, U32P), a, b, ia, ib, s)
        assert len(set(raw.tolist())) > n // 2
        assert all(O.orc_wrandd(int(u)) == d[k] for k, u in enumerate(raw))
        assert all(O.orc_wrandr(a, b, int(u)) == r[k] for k, u in enumerate(raw))
        assert all(O.orc_wrandi(ia, ib, s, int(u)) == iv[k] for k, u in enumerate(raw))


def test_multimask_against_reference():
    """makeMultiMask on break lists of 0-6 entries, unsorted, with repeats and
    breaks at bit 0 (GeneUtils.cpp:156-175)."""
    R, O = ref(), orc()
    rng = np.random.default_rng(7)
    cases = [[], [0], [62], [0, 0], [10, 4], [4, 10], [0, 32, 2]]
    for _ in range(3000):
        k = int(rng.integers(0, 7))
        cases.append((2 * rng.integers(0, 32, k)).tolist())
    for br in cases:
        a = np.array(br, np.uint32) if br else np.zeros(1, np.uint32)
        assert R.ref_multimask(p(a, U32P), len(br)) == O.orc_multimask(p(a, U32P), len(br)), br


@pytest.mark.parametrize("genome_size", [32, 100, 256, 1000])
def test_crossover_against_reference(genome_size):
    """GeneUtils::crossOver with 0-9 breaks (GeneUtils.cpp:184-246) against the
    oracle fed the breaks the reference drew."""
    R, O = ref(), orc()
    nb = (genome_size + 31) // 32
    nbits = nb * 64
    rng = np.random.default_rng(genome_size)
    for t, st in enumerate(states(genome_size, 60)):
        C_ = t % 10
        parent = rng.integers(0, 2**63, 2 * nb, dtype=np.uint64) * 2 + rng.integers(0, 2, 2 * nb, dtype=np.uint64)
        out_r = np.zeros(2 * nb, np.uint64)
        raw = np.zeros(max(C_, 1), np.uint32)
        R.ref_crossover(p(out_r, U64P), p(parent, U64P), genome_size, C_, p(st, U32P), p(raw, U32P))
        br = np.array([O.orc_wrandi(0, nbits, 2, int(u)) for u in raw[:C_]] or [0], np.uint32)
        out_o = np.zeros(2 * nb, np.uint64)
        O.orc_crossover(p(out_o, U64P), p(parent, U64P), genome_size, p(br, U32P), C_)
        assert np.array_equal(out_r, out_o), (genome_size, C_, br[:C_])


@pytest.mark.parametrize("nblocks", [1, 8, 31])
def test_free_reco_against_reference(nblocks):
    """GeneUtils::freeReco (GeneUtils.cpp:322-360)."""
    R, O = ref(), orc()
    rng = np.random.default_rng(nblocks)
    for st in states(nblocks + 100, 20):
        parent = rng.integers(0, 2**63, 2 * nblocks, dtype=np.uint64) * 2 + 1
        out_r = np.zeros(2 * nblocks, np.uint64)
        raw = np.zeros(2 * nblocks, np.uint32)
        R.ref_free_reco(p(out_r, U64P), p(parent, U64P), nblocks, p(st, U32P), p(raw, U32P))
        out_o = np.zeros(2 * nblocks, np.uint64)
        O.orc_free_reco(p(out_o, U64P), p(parent, U64P), nblocks, p(raw, U32P))
        assert np.array_equal(out_r, out_o)


@pytest.mark.parametrize("genome_size", [32, 256, 512])
def test_mutate_nucs_against_reference(genome_size):
    """GeneUtils::mutateNucs with 1-12 mutations (GeneUtils.cpp:112-141)."""
    R, O = ref(), orc()
    nb = (2 * genome_size + 63) // 64
    rng = np.random.default_rng(genome_size + 5)
    for t, st in enumerate(states(genome_size + 9, 40)):
        k = 1 + t % 12
        g0 = rng.integers(0, 2**63, nb, dtype=np.uint64)
        g_r = g0.copy()
        raw = np.zeros(2 * k, np.uint32)
        R.ref_mutate_nucs(p(g_r, U64P), genome_size, k, p(st, U32P), p(raw, U32P))
        g_o = g0.copy()
        pos = np.ascontiguousarray(raw[0::2])
        msk = np.ascontiguousarray(raw[1::2])
        O.orc_mutate_nucs(p(g_o, U64P), genome_size, k, p(pos, U32P), p(msk, U32P))
        assert np.array_equal(g_r, g_o)
        assert not np.array_equal(g_r, g0) or k % 2 == 0


@pytest.mark.parametrize("prob,n", [(1e-3, 512), (1e-3, 2000), (0.01, 200), (0.05, 64), (0.3, 40)])
def test_binomial_getn_against_reference(prob, n):
    """BinomialDist::create(p, n, 1e-6) + getN (BinomialDist.cpp:10-104) against
    the oracle's table on 200k draws and every table boundary +- 1 ulp."""
    R, O = ref(), orc()
    tab = np.zeros(64)
    nbins = O.orc_bino_table(prob, n, 1e-6, p(tab, F64P), 64)
    assert nbins > 0
    rng = np.random.default_rng(n)
    u = np.concatenate([rng.random(200000), tab[:nbins], np.nextafter(tab[:nbins], 0), np.nextafter(tab[:nbins], 2),
                        [0.0, 1.0 - 2**-53]])
    u = np.ascontiguousarray(u)
    got = np.zeros(len(u), np.int32)
    assert R.ref_bino_getn(prob, n, 1e-6, p(u, F64P), len(u), p(got, I32P)) == 0
    mine = np.array([O.orc_bino_getn(p(tab, F64P), nbins, float(x)) for x in u], np.int32)
    assert np.array_equal(got, mine)


def test_betai_against_reference():
    """bino_tools beta (the regularised incomplete beta, bino_tools.cpp:94)."""
    R, O = ref(), orc()
    rng = np.random.default_rng(3)
    for _ in range(2000):
        a = float(rng.uniform(0.5, 600))
        b = float(rng.uniform(0.5, 600))
        x = float(rng.random())
        assert R.ref_beta(a, b, x) == O.orc_betai(a, b, x), (a, b, x)
