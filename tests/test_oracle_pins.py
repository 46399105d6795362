"""Pins of oracle functions whose rules no invariant or statistic fixes:
Fertility's boundaries, RandomPair's k-th-unpaired rule under fixed draws,
MultiEvaluator's multiplicative mode (with SingleEvaluator's float cast and
cumulation), NPPCapacity's coastal / water / Oceania-Miami branches and the
deterministic exponential they use.  Every expected value below is worked out
by hand from the cited reference lines and written as a literal, so that a
slip in the oracle (a non-strict comparison, a reversed scan, a dropped term
or cast) fails here.  CPU only."""
import math

import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd.grids import rect_grid
from oracle.oracle import Oracle, lib

ALIVE, FERTILE = abi.QHG_LIFE_ALIVE, abi.QHG_LIFE_FERTILE


def _agents(gender, age, last_birth, life, cell=None):
    n = len(gender)
    return {"id": np.arange(1, n + 1, dtype=np.int64),
            "cell": np.zeros(n, dtype=np.int32) if cell is None else np.asarray(cell, dtype=np.int32),
            "life": np.asarray(life, dtype=np.uint8), "gender": np.asarray(gender, dtype=np.uint8),
            "birth_time": -np.asarray(age, dtype=np.float64), "age": np.asarray(age, dtype=np.float64),
            "last_birth": np.asarray(last_birth, dtype=np.float64),
            "mate_id": np.full(n, -3, dtype=np.int64), "num_babies": np.zeros(n, dtype=np.int32)}


def _empty_agents():
    return _agents([], [], [], [])


# -- Fertility (QHG4/actions/Fertility.cpp:49-74) -----------------------------
# Female: FERTILE iff age > min && age < max && (t - last_birth) > interbirth;
# male: FERTILE iff age > min; otherwise ALIVE.  All comparisons strict.
# Step 0 runs at t = 0.  min 15, max 45, interbirth 3.
FERTILITY_CASES = [
    # gender, age, last_birth, life before, life after
    (0, 15.0, -10.0, FERTILE, ALIVE),      # age == min: not above it
    (0, 15.5, -5.0, ALIVE, FERTILE),       # inside every bound
    (0, 45.0, -10.0, FERTILE, ALIVE),      # age == max: not below it
    (0, 44.9, -3.0, FERTILE, ALIVE),       # t - last_birth == interbirth
    (0, 44.9, -3.0001, ALIVE, FERTILE),    # just past the interbirth time
    (0, 50.0, -10.0, ALIVE, ALIVE),        # too old (females only)
    (0, 10.0, -10.0, FERTILE, ALIVE),      # too young
    (1, 15.0, -10.0, FERTILE, ALIVE),      # male at min: not above it
    (1, 15.0001, -10.0, ALIVE, FERTILE),   # male just above min
    (1, 50.0, -10.0, ALIVE, FERTILE),      # males have no upper bound
    (1, 20.0, 0.0, ALIVE, FERTILE),        # males ignore the interbirth time
]


def fertility_scenario():
    g, a, lb, before, _ = zip(*FERTILITY_CASES)
    params = scenarios.make_params(5, [(abi.ACT_FERTILITY, 1)], fert_min_age=15.0, fert_max_age=45.0,
                                   fert_interbirth=3.0)
    return {"name": "fertility", "grid": rect_grid(2, 2), "env": {}, "params": params,
            "agents": _agents(g, a, lb, before)}


def test_fertility_truth_table():
    o = Oracle(fertility_scenario())
    o.pre_loop()
    o.step()
    got = o.agents()["life"]
    want = np.array([c[4] for c in FERTILITY_CASES], dtype=np.uint8)
    assert np.array_equal(got, want), (got, want)


# -- RandomPair's rule (QHG4/actions/RandomPair.cpp:212-235, :249-272) ---------
def _pair_cell(nchoose, ncand, draws):
    import ctypes as C
    u = np.asarray(draws, dtype=np.uint32)
    taken = np.zeros(ncand, dtype=np.uint8)
    pick = np.full(nchoose, -1, dtype=np.int32)
    L = lib()
    L.orc_pair_cell.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_uint32), C.POINTER(C.c_uint8),
                                C.POINTER(C.c_int32)]
    L.orc_pair_cell(nchoose, ncand, abi.ptr(u, C.c_uint32), abi.ptr(taken, C.c_uint8), abi.ptr(pick, C.c_int32))
    return pick.tolist()


@pytest.mark.parametrize("nchoose,ncand,draws,want", [
    # left 4: k = floor(4 * 0.75) = 3 -> candidate 3; left 3: k = 0 -> candidate 0;
    # left 2: k = floor(2 * (2^32-1)/2^32) = 1 -> second unpaired of {1, 2} = 2
    (3, 4, [0xC0000000, 0x00000000, 0xFFFFFFFF], [3, 0, 2]),
    # left 5: k = floor(2.5) = 2 -> candidate 2; left 4: k = floor(4 * 0.875) = 3 ->
    # fourth unpaired of {0, 1, 3, 4} = 4
    (2, 5, [0x80000000, 0xE0000000], [2, 4]),
    # left 3: k = floor(3 * 0.5) = 1 -> 1; left 2: k = floor(2 * 0.5) = 1 -> second of
    # {0, 2} = 2; left 1: k = 0 -> 0
    (3, 3, [0x80000000, 0x80000000, 0x80000000], [1, 2, 0]),
    # one chooser, one candidate: k = 0 whatever the draw
    (1, 1, [0xFFFFFFFF], [0]),
])
def test_pair_rule_fixed_draws(nchoose, ncand, draws, want):
    assert _pair_cell(nchoose, ncand, draws) == want


def _pair_scenario(genders, seed):
    n = len(genders)
    params = scenarios.make_params(seed, [(abi.ACT_RANDOMPAIR, 1)])
    return {"name": "pair", "grid": rect_grid(2, 2), "env": {}, "params": params,
            "agents": _agents(genders, [20.0] * n, [-10.0] * n, [FERTILE] * n)}


@pytest.mark.parametrize("genders", [[1, 0, 1, 0, 1], [0, 1, 0, 0, 1, 0]])
def test_pairing_step_uses_the_rule(genders):
    """A whole step of the oracle pairs exactly as the rule above dictates:
    the sex with fewer candidates chooses (females when F <= M), choosers and
    candidates in id order, each chooser with its own pairing draw."""
    seed = 77
    o = Oracle(_pair_scenario(genders, seed))
    o.pre_loop()
    o.step()
    a = o.agents()
    ids = a["id"]
    fem = [int(i) for i, g in zip(ids, a["gender"]) if g == 0]
    mal = [int(i) for i, g in zip(ids, a["gender"]) if g == 1]
    chooser, cand = (fem, mal) if len(fem) <= len(mal) else (mal, fem)
    draws = [lib().orc_draw(seed, 0, abi.STREAM_PAIR, i, 0) for i in chooser]
    pick = _pair_cell(len(chooser), len(cand), draws)
    mate = dict(zip(ids.tolist(), a["mate_id"].tolist()))
    for ch, k in zip(chooser, pick):
        assert mate[ch] == cand[k] and mate[cand[k]] == ch
    unpaired = set(cand) - {cand[k] for k in pick}
    assert all(mate[i] == -3 for i in unpaired)


# -- MultiEvaluator, multiplicative mode ---------------------------------------
# MultiEvaluator::multiplySingleWeights (QHG4/actions/MultiEvaluator.cpp:308-342):
# W = 1, then W *= single_e * weight_e for each evaluator, then an in-row
# prefix sum.  SingleEvaluator::calcValues (SingleEvaluator.cpp:174-207):
# s = PolyLine((float)input) or the raw input, clamped at 0; then
# exchangeAndCumulate (:216-243): row [s_c, s_n0 .. s_n3] with -1 neighbours
# giving 0, cumulative when bCumulate.
#
# 3x3 lattice, cell = 3y + x, neighbour slots (+x, +y, -x, -y).
# Evaluator 0: altitude through PolyLine (2000, 0) - (3000, 1), weight 2.
#   Cell 0's altitude 2500.0001 is 2500 as a float, so s = 0.5 exactly
#   (without the cast it would be 0.5000001).
# Evaluator 1: base NPP as is, weight 0.5; cell 8's -1 clamps to 0.
ALT = [2500.0001, 2250.0, 2000.0, 2750.0, 3000.0, 2100.0, -100.0, 4000.0, 2600.0]
NPP = [1.0, 2.0, 0.5, 1.5, 4.0, 3.0, 1.0, 0.25, -1.0]
# s0 = [0.5, 0.25, 0, 0.75, 1, 0.1, 0, 1, 0.6]; s1 = [1, 2, 0.5, 1.5, 4, 3, 1, 0.25, 0]
# per-cell product p = s0 * s1 (weights 2 and 0.5 cancel) =
#   [0.5, 0.5, 0, 1.125, 4, 0.3, 0, 0.25, 0]
ROWS_PLAIN = {
    0: [0.5, 1.0, 2.125, 2.125, 2.125],     # p0, p1, p3, hole, hole
    4: [4.0, 4.3, 4.55, 5.675, 6.175],      # p4, p5, p7, p3, p1
    8: [0.0, 0.0, 0.0, 0.25, 0.55],         # p8, hole, hole, p7, p5
}
# evaluator 0 cumulating: its row is the running sum of s0 over (c, n0..n3),
# each entry then multiplied by s1 of the same slot, then the prefix sum:
#   cell 0: s0 run [0.5, 0.75, 1.5, 1.5, 1.5] x s1 [1, 2, 1.5, 0, 0] = [0.5, 1.5, 2.25, 0, 0]
#   cell 4: s0 run [1, 1.1, 2.1, 2.85, 3.1] x s1 [4, 3, 0.25, 1.5, 2] = [4, 3.3, 0.525, 4.275, 6.2]
#   cell 8: s0 run [0.6, 0.6, 0.6, 1.6, 1.7] x s1 [0, 0, 0, 0.25, 3] = [0, 0, 0, 0.4, 5.1]
ROWS_CUMULATE = {
    0: [0.5, 2.0, 4.25, 4.25, 4.25],
    4: [4.0, 7.3, 7.825, 12.1, 18.3],
    8: [0.0, 0.0, 0.0, 0.4, 5.5],
}


def mul_eval_scenario(cumulate0):
    params = scenarios.make_params(
        3, [(abi.ACT_MULTIEVAL, 1, "MultiEvaluator[Alt+NPP]")],
        me_mode=abi.ME_MUL_SIMPLE, me_num=2, me_input=[abi.EV_ALTITUDE, abi.EV_BASE_NPP],
        me_cumulate=[int(cumulate0), 0], me_weight=[2.0, 0.5], me_pl=[[(2000.0, 0.0), (3000.0, 1.0)], []])
    env = {"altitude": np.array(ALT), "base_npp": np.array(NPP)}
    return {"name": "mul-eval", "grid": rect_grid(3, 3), "env": env, "params": params, "agents": _empty_agents()}


@pytest.mark.parametrize("cumulate0,rows", [(False, ROWS_PLAIN), (True, ROWS_CUMULATE)])
def test_multievaluator_mul_rows(cumulate0, rows):
    o = Oracle(mul_eval_scenario(cumulate0))
    o.pre_loop()
    o.step()
    w = o.env_weights().reshape(9, 5)
    for c, want in rows.items():
        np.testing.assert_allclose(w[c], want, rtol=1e-14, atol=0, err_msg="cell %d" % c)


# -- NPPCapacity (QHG4/actions/NPPCapacity.cpp:138-217) ------------------------
# K_min 5, K_max 60, NPP_min 0.1, NPP_max 1.5, coastal factor 0.1 within
# latitudes (-30, 30), water factor 0.05, efficiency 0.8.  Ramp:
# K = 5 + NPP * 60 / 1.4; coastal adds 6, water w adds w * 3; cap 60; x 0.8.
# Oceania (lon in (115, 150), lat in (-12, 1)) with NPP < NPP_min takes the
# Miami NPP 0.000475 * min(3000 / (1 + e^(1.315 - 0.119 T)),
# 3000 (1 - e^(-0.000664 P))) (QHG4/core/NPPCalcMiami.cpp:28-41).
T_750 = (1.315 - math.log(3.0)) / 0.119      # 3000 / (1 + 3) = 750
P_300 = -math.log(0.9) / 0.000664            # 3000 * (1 - 0.9) = 300
NPP_CELLS = [
    # alt, npp, lon, lat, coastal, water, temp, rain, K
    (-10.0, 0.75, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 0.0),          # sea: 0
    (500.0, 0.75, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 208.0 / 7),    # (5 + 225/7) * 0.8
    (2000.0, 1.0, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 148.0 / 7),    # altitude factor 1/2: (5 + 150/7) * 0.8
    (3000.0, 1.0, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 4.0),          # factor 0: NPP 0 < min -> K_min * 0.8
    (500.0, 0.05, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 4.0),          # NPP < min outside Oceania -> K_min
    (500.0, 0.75, 10.0, 10.0, 1, 0.0, 20.0, 1000.0, 241.6 / 7),    # coastal inside the band: (260/7 + 6) * 0.8
    (500.0, 0.75, 10.0, 40.0, 1, 0.0, 20.0, 1000.0, 208.0 / 7),    # coastal outside the band: no bonus
    (500.0, 0.75, 10.0, 40.0, 0, 0.5, 20.0, 1000.0, 1082.0 / 35),  # water 0.5: (260/7 + 1.5) * 0.8
    (500.0, 1.4, 10.0, 10.0, 1, 0.0, 20.0, 1000.0, 48.0),          # 65 + 6 capped at 60, * 0.8
    (500.0, 2.0, 10.0, 40.0, 0, 0.0, 20.0, 1000.0, 48.0),          # NPP > max -> K_max
    (500.0, 0.05, 130.0, -5.0, 0, 0.0, 1000.0, P_300, 311.0 / 35),  # Miami min(3000, 300) -> 0.1425
    (500.0, 0.05, 130.0, -5.0, 0, 0.0, T_750, 1e7, 227.0 / 14),    # Miami min(750, 3000) -> 0.35625
    (500.0, 0.5, 130.0, -5.0, 0, 0.0, T_750, 1e7, 148.0 / 7),      # Oceania but NPP >= min: kept
    (500.0, 0.05, 115.0, -5.0, 0, 0.0, T_750, 1e7, 4.0),           # lon == 115 is outside the box
]


def npp_scenario():
    cols = list(zip(*NPP_CELLS))
    env = {"altitude": np.array(cols[0]), "base_npp": np.array(cols[1]), "longitude": np.array(cols[2]),
           "latitude": np.array(cols[3]), "coastal": np.array(cols[4], dtype=np.uint8),
           "water": np.array(cols[5]), "temperature": np.array(cols[6]), "rainfall": np.array(cols[7])}
    params = scenarios.make_params(
        4, [(abi.ACT_NPPCAP, 1)], npp_min=0.1, npp_max=1.5, npp_k_min=5.0, npp_k_max=60.0,
        npp_coastal_factor=0.1, npp_coastal_min_lat=-30.0, npp_coastal_max_lat=30.0, npp_water_factor=0.05,
        npp_efficiency=0.8)
    return {"name": "npp", "grid": rect_grid(len(NPP_CELLS), 1), "env": env, "params": params,
            "agents": _empty_agents()}


def test_npp_capacity_branches():
    o = Oracle(npp_scenario())
    o.pre_loop()
    want = np.array([c[8] for c in NPP_CELLS])
    np.testing.assert_allclose(o.capacities(), want, rtol=1e-13, atol=0)


# -- the deterministic exponential --------------------------------------------
def test_exp_against_libm():
    import ctypes as C
    L = lib()
    L.orc_exp.argtypes = [C.c_double]
    L.orc_exp.restype = C.c_double
    assert L.orc_exp(0.0) == 1.0
    assert L.orc_exp(-1000.0) == 0.0 and L.orc_exp(1000.0) == math.inf
    assert math.isnan(L.orc_exp(math.nan))
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-20, 20, 20000), rng.uniform(-700, 700, 5000),
                         np.linspace(-0.35, 0.35, 2001), [math.log(2.0), -math.log(2.0), 1.0, -1.0]])
    got = np.array([L.orc_exp(float(x)) for x in xs])
    ref = np.exp(xs)
    ulp = np.abs(got - ref) / np.spacing(ref)
    assert ulp.max() <= 2.0, ulp.max()
    # e^a e^b = e^(a+b) within rounding; e^1 is e
    assert abs(L.orc_exp(1.0) - math.e) <= 2 * np.spacing(math.e)
