"""AssortativePairHamming in the oracle: the genome distance and the RankTable
pairing pinned against the reference's own sources (compiled into
oracle/_ref/libqhg_ref.so from QHG4/genes/GeneUtils.cpp and
QHG4/utils/RankTable.cpp), and against what the algorithm fixes by itself:
identical genomes are at distance 0, swapped strands too, a table of distinct
values pairs greedily by maximum, a cutoff above every value pairs nobody.
"""
import ctypes as C
import os

import numpy as np
import pytest

from oracle.oracle import Oracle, lib as olib
from qhg4_amd import abi, scenarios

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
REF = os.path.join(ROOT, "oracle", "_ref", "libqhg_ref.so")
need_ref = pytest.mark.skipif(not os.path.exists(REF),
                              reason="reference utilities not built (make -C oracle ref needs the reference tree)")

U32P = C.POINTER(C.c_uint32)
U64P = C.POINTER(C.c_uint64)
I32P = C.POINTER(C.c_int32)
F32P = C.POINTER(C.c_float)


def orc():
    L = olib()
    L.orc_calc_dist.argtypes = [U64P, U64P, C.c_int32]
    L.orc_calc_dist.restype = C.c_int32
    L.orc_rank_pairs.argtypes = [C.c_int32, C.c_int32, F32P, C.c_float, C.c_int32, U32P, C.c_int32, I32P, I32P]
    L.orc_rank_pairs.restype = C.c_int32
    return L


def ref():
    L = C.CDLL(REF)
    L.ref_calc_dist.argtypes = [U64P, U64P, C.c_int]
    L.ref_calc_dist.restype = C.c_int
    L.ref_rank_pairs.argtypes = [C.c_int, C.c_int, F32P, C.c_float, C.c_int, U32P, I32P, U32P, C.c_int,
                                 C.POINTER(C.c_int)]
    L.ref_rank_pairs.restype = C.c_int
    return L


def p(a, t):
    return a.ctypes.data_as(t)


def dist(g1, g2, arg):
    return orc().orc_calc_dist(p(g1, U64P), p(g2, U64P), arg)


def rank_pairs(table, cutoff, permute, u):
    nf, nm = table.shape
    t = np.ascontiguousarray(table, np.float32).copy()
    u = np.ascontiguousarray(u, np.uint32)
    pairs = np.zeros(2 * max(1, min(nf, nm)), np.int32)
    used = C.c_int32(0)
    n = orc().orc_rank_pairs(nf, nm, p(t, F32P), cutoff, permute, p(u, U32P), len(u), p(pairs, I32P),
                             C.byref(used))
    return [tuple(pairs[2 * k:2 * k + 2]) for k in range(max(n, 0))], n, used.value


# ---------------------------------------------------------------- distance

def test_distance_closed_forms():
    rng = np.random.default_rng(1)
    w = 8
    g = rng.integers(0, 2 ** 63, 2 * w, dtype=np.int64).astype(np.uint64)
    assert dist(g, g.copy(), 256) == 0
    # one nucleotide differs (any of the three non-zero 2-bit changes): distance 1
    for v in (1, 2, 3):
        h = g.copy()
        h[3] ^= np.uint64(v) << np.uint64(10)
        assert dist(g, h, 256) == 1
    # the second parent's strands swapped: the crossed comparison finds 0
    h = np.concatenate([g[w:], g[:w]])
    assert dist(g, h, 256) == 0
    # every nucleotide of strand 0 differs, strand 1 equal: 32 per word
    h = g.copy()
    h[:w] ^= np.uint64(0x5555555555555555)
    assert dist(g, h, 256) == 32 * w
    # size_arg 8 (the block count AssortativePairHam passes) covers one word
    # per strand position: words 0 and 1 only
    h = g.copy()
    h[2:] ^= np.uint64(0xFFFFFFFFFFFFFFFF)
    assert dist(g, h, 8) == 0
    h[1] ^= np.uint64(3)
    assert dist(g, h, 8) == 1


@need_ref
@pytest.mark.parametrize("size", [20, 64, 100, 256, 1000, 1100, 2048])
def test_distance_matches_reference(size):
    R = ref()
    rng = np.random.default_rng(size)
    w = (size + 31) // 32
    for trial in range(40):
        g1 = rng.integers(0, 2 ** 63, 2 * w, dtype=np.int64).astype(np.uint64)
        g2 = g1.copy()
        k = rng.integers(0, 4 * w)
        idx = rng.integers(0, 2 * w, k)
        g2[idx] ^= rng.integers(0, 2 ** 63, k, dtype=np.int64).astype(np.uint64)
        if trial % 3 == 0:
            g2 = np.concatenate([g2[w:], g2[:w]])
        for arg in (w, size):   # the reference's call (block count) and a genome size
            assert dist(g1, g2, arg) == R.ref_calc_dist(p(g1, U64P), p(g2, U64P), arg), (size, trial, arg)


# ---------------------------------------------------------------- RankTable

def test_rank_pairs_greedy_on_distinct_values():
    """Distinct values: each round's maximum is unique, so the pairs are the
    greedy maximum matching whatever the draws (brute force alongside)."""
    rng = np.random.default_rng(5)
    for trial in range(30):
        nf, nm = rng.integers(1, 9, 2)
        t = rng.permutation(nf * nm).reshape(nf, nm).astype(np.float32) / (nf * nm) + np.float32(0.01)
        got, n, used = rank_pairs(t, 0.0, trial % 2, rng.integers(0, 2 ** 32, 4096, dtype=np.uint64))
        exp, tt = [], t.copy()
        while tt.max() > 0:
            i, j = np.unravel_index(np.argmax(tt), tt.shape)
            exp.append((i, j))
            tt[i, :] = 0
            tt[:, j] = 0
        assert got == exp, trial


def test_rank_pairs_ties_and_cutoff():
    rng = np.random.default_rng(6)
    t = np.ones((5, 7), np.float32)
    got, n, _ = rank_pairs(t, 0.5, 1, rng.integers(0, 2 ** 32, 4096, dtype=np.uint64))
    assert n == 5 and len({f for f, _ in got}) == 5 and len({m for _, m in got}) == 5
    got, n, used = rank_pairs(t, 1.5, 1, rng.integers(0, 2 ** 32, 16, dtype=np.uint64))
    assert n == 0 and used == 0
    # values below the cutoff never pair, values at it do
    t = np.array([[0.5, 0.2], [0.3, 0.1]], np.float32)
    got, n, _ = rank_pairs(t, 0.3, 0, np.zeros(8, np.uint32))
    assert got == [(0, 0)]


@need_ref
@pytest.mark.parametrize("permute", [0, 1])
def test_rank_pairs_match_reference(permute):
    """Same table, same WELL512 words: the same pairs in the same order and
    the same number of draws as the reference's RankTable."""
    R = ref()
    rng = np.random.default_rng(10 + permute)
    levels = np.array([0.0, 0.25, 0.5, 0.75, 0.875, 1.0], np.float32)
    for trial in range(300):
        nf, nm = (int(v) for v in rng.integers(1, 13, 2))
        if trial % 3 == 0:
            t = levels[rng.integers(0, len(levels), (nf, nm))]
        else:
            t = rng.random((nf, nm), dtype=np.float32) * np.float32(1.2) - np.float32(0.1)
        cutoff = float(np.float32([0.0, 0.3, 0.6, -0.2, 0.9][trial % 5]))
        state = rng.integers(0, 2 ** 32, 16, dtype=np.uint64).astype(np.uint32)
        pairs = np.zeros(2 * 13, np.int32)
        raw = np.zeros(4096, np.uint32)
        nraw = C.c_int(0)
        tt = np.ascontiguousarray(t, np.float32).copy()
        n = R.ref_rank_pairs(nf, nm, p(tt, F32P), cutoff, permute, p(state, U32P), p(pairs, I32P), p(raw, U32P),
                             len(raw), C.byref(nraw))
        assert 0 <= nraw.value <= len(raw)
        exp = [tuple(pairs[2 * k:2 * k + 2]) for k in range(n)]
        got, gn, used = rank_pairs(t, cutoff, permute, raw[:nraw.value])
        assert (got, used) == (exp, nraw.value), (trial, nf, nm, cutoff)


# ---------------------------------------------------------------- the action

def test_assort_pairs_within_cells():
    """OoANavGen2bitAssHamPop's pairing in the oracle: mates share a cell,
    are of opposite sex, and pair both ways (no moves, so the cells at the
    end of a step are those of its pairing)."""
    sc = scenarios.c2(seed=3, level=4, n_agents=3000, patch_cells=150, genetics=True, assort=(0.9, 1))
    sc["params"].wmove_prob = 0.0
    o = Oracle(sc)
    o.pre_loop()
    o.run(30)
    a = o.agents()
    idx = {int(i): k for k, i in enumerate(a["id"])}
    paired = [k for k in range(len(a["id"])) if a["mate_id"][k] >= 0]
    assert len(paired) > 100
    for k in paired:
        m = idx[int(a["mate_id"][k])]
        assert a["mate_id"][m] == a["id"][k]
        assert a["cell"][m] == a["cell"][k]
        assert a["gender"][m] != a["gender"][k]


def test_assort_cutoff_above_one_pairs_nobody():
    """Match values are at most 1: a cutoff above it leaves everybody
    unpaired, so no births."""
    sc = scenarios.c2(seed=4, level=4, n_agents=2000, patch_cells=120, genetics=True, assort=(1.01, 0))
    o = Oracle(sc)
    o.pre_loop()
    o.run(20)
    st = o.stats()
    assert st["births_total"] == 0
    assert (o.agents()["mate_id"] < 0).all()
