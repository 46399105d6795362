"""Pins for the oracle's building blocks (CPU only).

Each check compares the oracle against something it does not share code
with: published known-answer vectors, scipy, closed forms, brute force.
"""
import ctypes as C
import itertools
import math

import numpy as np
import pytest
import scipy.stats

from oracle.oracle import lib

P = C.POINTER


# --------------------------------------------------------------------------
# Philox4x32-10: Random123 known-answer vectors (kat_vectors, philox4x32 10)
# --------------------------------------------------------------------------
KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8]),
    ([0xFFFFFFFF] * 4, [0xFFFFFFFF] * 2, [0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD]),
    ([0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344], [0xA4093822, 0x299F31D0],
     [0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1]),
]


@pytest.mark.parametrize("ctr,key,expect", KAT)
def test_philox_kat(ctr, key, expect):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    assert list(o) == expect


def test_draw_counter_layout():
    """draw(seed, step, stream, id, n) = word n&3 of Philox(ctr={step,
    stream<<16 | n>>2, id_lo, id_hi}, key=seed)."""
    seed, step, stream, aid = 0x0123456789ABCDEF, 77, 5, 0x1_0000_0007
    for n in range(9):
        c = (C.c_uint32 * 4)(step, (stream << 16) | (n >> 2), aid & 0xFFFFFFFF, aid >> 32)
        k = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, seed >> 32)
        o = (C.c_uint32 * 4)()
        lib().orc_philox4x32_10(c, k, o)
        assert lib().orc_draw(seed, step, stream, aid, n) == o[n & 3]


def test_draws_look_uniform():
    """Philox words keyed by consecutive ids: KS against U(0,1)."""
    u = np.array([lib().orc_draw(42, 3, 2, i, 0) for i in range(20000)], dtype=np.float64) / 2.0 ** 32
    assert scipy.stats.kstest(u, "uniform").pvalue > 1e-3


# --------------------------------------------------------------------------
# WELL512 conversions (QHG4/utils/WELL512.h:33-39) on fixed 32-bit inputs
# --------------------------------------------------------------------------
EDGE = [0, 1, 2, 3, 0x7FFFFFFF, 0x80000000, 0xFFFFFFFE, 0xFFFFFFFF, 123456789]


@pytest.mark.parametrize("u", EDGE)
def test_wrandd(u):
    assert lib().orc_wrandd(u) == u / 4294967296.0
    assert 0.0 <= lib().orc_wrandd(u) < 1.0


@pytest.mark.parametrize("u", EDGE)
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 64, 1000])
def test_wrandr_pick_is_floor(u, n):
    # (int)wrandr(0, n) == floor(n*u / 2^32), exact in integers
    assert int(lib().orc_wrandr(0.0, float(n), u)) == (n * u) >> 32


@pytest.mark.parametrize("u", EDGE)
def test_wrandi(u):
    L = lib()
    assert L.orc_wrandi(0, 512, 2, u) == 2 * ((512 * u) >> 33)   # crossover break, even bit
    assert L.orc_wrandi(0, 1024, 2, u) == 2 * (u >> 23)          # mutation position
    assert L.orc_wrandi(1, 4, 1, u) == 1 + ((3 * u) >> 32)       # mutation mask 1..3
    assert L.orc_wrandi(0, 7, 1, u) == (7 * u) >> 32
    # s=2 with an odd range: r += r % s rounds the range up
    assert L.orc_wrandi(0, 5, 2, u) == 2 * ((6 * u) >> 33)
    assert 1 <= L.orc_wrandi(1, 4, 1, u) <= 3


# --------------------------------------------------------------------------
# makeMultiMask (QHG4/genes/GeneUtils.cpp:156-175)
# --------------------------------------------------------------------------
def multimask_closed_form(breaks):
    """Closed form of makeMultiMask.  The routine walks bits from the MSB and
    consumes breaks from the back of the vector, one per bit position: the
    breaks that take effect are the longest run at the back of the vector
    whose positions are non-zero and strictly decreasing (read backwards).
    Bit b then holds 1 - n%2 flipped once for every effective break > b."""
    n = len(breaks)
    eff = []
    prev = 65
    for v in reversed(breaks):
        if 1 <= v < prev:
            eff.append(v)
            prev = v
        else:
            break
    top = 1 - n % 2
    out = 0
    for b in range(64):
        flips = sum(1 for v in eff if v > b)
        out |= ((top ^ (flips & 1)) & 1) << b
    return out


def mm(breaks):
    arr = (C.c_uint32 * max(1, len(breaks)))(*breaks)
    return lib().orc_multimask(arr, len(breaks))


def test_multimask_sorted_textbook():
    # sorted breaks: bit b = 1 iff the number of breaks <= b is even
    for breaks in ([10], [2, 40], [0, 16, 62], [6, 8, 30, 50]):
        want = sum(1 << b for b in range(64) if sum(1 for v in breaks if v <= b) % 2 == 0)
        assert mm(breaks) == want, breaks


def test_multimask_bruteforce_unsorted():
    """All multisets of <= 3 even break positions in one block, in every
    draw order (covers unsorted vectors, repeats and a break at bit 0)."""
    pos = list(range(0, 64, 2))
    count = 0
    for n in range(0, 4):
        for combo in itertools.product(pos, repeat=n):
            assert mm(list(combo)) == multimask_closed_form(list(combo)), combo
            count += 1
    assert count == 1 + 32 + 32 ** 2 + 32 ** 3


# --------------------------------------------------------------------------
# crossOver / mutateNucs invariants
# --------------------------------------------------------------------------
def crossover(parent, genome_size, breaks):
    nb = (genome_size + 31) // 32
    out = (C.c_uint64 * (2 * nb))()
    par = (C.c_uint64 * (2 * nb))(*parent)
    br = (C.c_uint32 * max(1, len(breaks)))(*breaks)
    lib().orc_crossover(out, par, genome_size, br, len(breaks))
    return list(out)


def test_crossover_no_breaks_copies():
    rng = np.random.default_rng(1)
    par = [int(x) for x in rng.integers(0, 2 ** 63, 16)]
    assert crossover(par, 256, []) == par


def test_crossover_single_break_swaps_tails():
    """One break at bit position p (whole genome of 512 bits per strand):
    strand 0 keeps strand 0 below p and takes strand 1 from p on."""
    rng = np.random.default_rng(2)
    par = [int(x) for x in rng.integers(0, 2 ** 63, 16)]
    g0 = sum(par[i] << (64 * i) for i in range(8))
    g1 = sum(par[8 + i] << (64 * i) for i in range(8))
    for p in (2, 64, 130, 300, 510):
        out = crossover(par, 256, [p])
        o0 = sum(out[i] << (64 * i) for i in range(8))
        o1 = sum(out[8 + i] << (64 * i) for i in range(8))
        low = (1 << p) - 1
        full = (1 << 512) - 1
        assert o0 == (g0 & low) | (g1 & ~low & full), p
        assert o1 == (g1 & low) | (g0 & ~low & full), p


def test_crossover_loci_come_from_one_parent_strand():
    """Every output bit comes from one of the two parent strands at the same
    position, and the two outputs are complementary selections."""
    rng = np.random.default_rng(3)
    for _ in range(200):
        par = [int(x) for x in rng.integers(0, 2 ** 63, 16)]
        nbk = int(rng.integers(1, 6))
        br = [int(2 * rng.integers(0, 256)) for _ in range(nbk)]
        out = crossover(par, 256, br)
        for i in range(8):
            a, b, x, y = par[i], par[8 + i], out[i], out[8 + i]
            # x takes each bit from a or b ...
            assert ((x ^ a) & (x ^ b)) == 0
            # ... and y takes the other one
            assert (x ^ y) == (a ^ b)


def test_mutate_flips_one_nucleotide():
    g = (C.c_uint64 * 16)()
    for pos_draw, mask_draw in ((0, 0), (0xFFFFFFFF, 0xFFFFFFFF), (0x12345678, 0x9ABCDEF0)):
        for i in range(16):
            g[i] = 0
        pd = (C.c_uint32 * 1)(pos_draw)
        md = (C.c_uint32 * 1)(mask_draw)
        lib().orc_mutate_nucs(g, 512, 1, pd, md)
        bits = [(i, g[i]) for i in range(16) if g[i]]
        assert len(bits) == 1
        blk, val = bits[0]
        pos = 2 * (pos_draw >> 23)
        assert blk == pos // 64 and val == ((1 + ((3 * mask_draw) >> 32)) << (pos % 64))


# --------------------------------------------------------------------------
# Binomial mutation-count table vs scipy (BinomialDist::buildTable)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("p,n", [(1e-3, 512), (1e-2, 512), (0.05, 100), (0.3, 40)])
def test_binomial_table_matches_scipy(p, n):
    cdf = (C.c_double * 64)()
    cnt = lib().orc_bino_table(p, n, 1e-6, cdf, 64)
    assert cnt > 0
    for k in range(cnt):
        assert abs(cdf[k] - scipy.stats.binom.cdf(k, n, p)) < 2e-6, (k, cdf[k])
    # the table stops once the tail is below eps
    assert 1 - cdf[cnt - 1] <= 1e-6 + 1e-9


def test_betai_closed_forms():
    L = lib()
    for x in (0.1, 0.37, 0.9):
        assert abs(L.orc_betai(1.0, 1.0, x) - x) < 1e-7           # I_x(1,1) = x
        assert abs(L.orc_betai(2.0, 1.0, x) - x * x) < 1e-7       # I_x(2,1) = x^2
        assert abs(L.orc_betai(1.0, 3.0, x) - (1 - (1 - x) ** 3)) < 1e-7


def test_getn_inverse_cdf():
    cdf = (C.c_double * 64)()
    cnt = lib().orc_bino_table(0.01, 512, 1e-6, cdf, 64)
    L = lib()
    assert L.orc_bino_getn(cdf, cnt, 0.0) == 0
    for k in range(cnt):
        assert L.orc_bino_getn(cdf, cnt, cdf[k]) == k            # u == cdf[k] stays at k
        assert L.orc_bino_getn(cdf, cnt, math.nextafter(cdf[k], 2)) == k + 1


# --------------------------------------------------------------------------
# PolyLine::getVal (QHG4/utils/PolyLine.cpp:60-89)
# --------------------------------------------------------------------------
def test_polyline():
    xs = [-0.1, 0.0, 0.1, 1500.0, 2000.0, 3000.0]
    vs = [0.0, 0.1, 0.01, 1.0, 1.0, -9999.0]
    X = (C.c_double * 6)(*xs)
    V = (C.c_double * 6)(*vs)
    f = lambda x: lib().orc_polyline(6, X, V, x)  # noqa: E731
    assert f(-5.0) == 0.0                     # clamps below
    assert f(4000.0) == -9999.0               # clamps above
    for x, v in zip(xs, vs):
        # passes through the points (interior ones via slope*dx: 1 ulp-level)
        assert abs(f(x) - v) <= 1e-12 * max(1.0, abs(v))
    assert abs(f(750.05) - (0.01 + (1.0 - 0.01) * (750.05 - 0.1) / 1499.9)) < 1e-12
    assert abs(f(2500.0) - (1.0 + (-10000.0) * 0.5)) < 1e-9
    ident = lambda x: lib().orc_polyline(0, X, V, x)  # noqa: E731
    assert ident(123.25) == 123.25            # no segments: identity
