"""Pins of the oracle's OldAgeDeath and ATanDeath (and its deterministic
arctangent) against what the mathematics fixes: libm atan to a few ulp,
exact symmetries, and the death laws the reference's execute() implements
(QHG4/actions/OldAgeDeath.cpp execute; QHG4/actions/ATanDeath.cpp:50-90):
  OldAgeDeath: dies iff age > M + U, U ~ uniform[1 - uM, 1 + uM)
  ATanDeath:   dies with probability 1/2 + s*atan(k*(age - M))/pi,
               s = (pi/2 - 0.001) / atan(k * range)
checked as frequencies over many agents of one age (binomial bounds) and at
the ages where the law is certain."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.oracle import Oracle, lib
from qhg4_amd import abi, scenarios
from qhg4_amd.grids import rect_grid


def orc_atan(x):
    L = lib()
    L.orc_atan.argtypes = [C.c_double]
    L.orc_atan.restype = C.c_double
    return L.orc_atan(x)


def ulp(x):
    return math.ulp(abs(x)) if x != 0 else math.ulp(0.0)


def test_atan_against_libm():
    rng = np.random.default_rng(1)
    xs = list(rng.uniform(-3, 3, 20000)) + list(rng.uniform(-200, 200, 5000))
    xs += list(10.0 ** rng.uniform(-12, 12, 5000)) + [k / 8 for k in range(9)] + [k / 16 for k in range(17)]
    xs += [0.0, 1.0, -1.0, 1e300, -1e300, 5e-324, 0.9999999999999999, 1.0000000000000002]
    worst = 0.0
    for x in xs:
        got, ref = orc_atan(float(x)), math.atan(float(x))
        err = abs(got - ref) / ulp(ref)
        worst = max(worst, err)
        assert err <= 3.0, (x, got, ref)
    assert worst <= 3.0


def test_atan_symmetries():
    rng = np.random.default_rng(2)
    for x in rng.uniform(-50, 50, 2000):
        assert orc_atan(-x) == -orc_atan(x)
    assert orc_atan(0.0) == 0.0
    assert orc_atan(1.0) == math.pi / 4
    assert orc_atan(1e308) == math.pi / 2


def population(ages, action, **kw):
    """One agent per entry of ages on a 64x64 lattice; only the death action."""
    n = len(ages)
    grid = rect_grid(64, 64, periodic=False)
    rng = np.random.default_rng(5)
    agents = {
        "id": np.arange(1, n + 1, dtype=np.int64),
        "cell": rng.integers(0, grid["num_cells"], n).astype(np.int32),
        "life": np.full(n, abi.QHG_LIFE_ALIVE, dtype=np.uint8),
        "gender": (rng.random(n) < 0.5).astype(np.uint8),
        "birth_time": -np.asarray(ages, dtype=np.float64),
        "age": np.array(ages, dtype=np.float64),
        "last_birth": np.full(n, -1.0),
        "mate_id": np.full(n, -3, dtype=np.int64),
        "num_babies": np.zeros(n, dtype=np.int32),
    }
    params = scenarios.make_params(9, [(action, 1)], **kw)
    return {"name": "deaths", "grid": grid, "env": {}, "params": params, "agents": agents}


def died(sc):
    o = Oracle(sc)
    o.pre_loop()
    o.step()
    alive = set(o.agents()["id"].tolist())
    ids = sc["agents"]["id"]
    return np.array([i not in alive for i in ids.tolist()])


def binom_ok(k, n, p, z=5.0):
    sd = math.sqrt(max(n * p * (1 - p), 1e-12))
    return abs(k - n * p) <= z * sd + 1


def test_oldagedeath_law():
    M, u = 60.0, 0.1          # U ~ [1 - 6, 1 + 6) = [-5, 7): certain below 55, certain above 67
    ages = np.repeat([40.0, 54.9, 57.0, 61.0, 64.0, 67.1, 90.0], 8000)
    d = died(population(ages, abi.ACT_OLDAGEDEATH, oad_max_age=M, oad_uncertainty=u))
    for a in np.unique(ages):
        m = ages == a
        p = min(1.0, max(0.0, (a - M - (1 - u * M)) / (2 * u * M)))
        assert binom_ok(int(d[m].sum()), int(m.sum()), p), (a, d[m].mean(), p)
    assert not d[ages < 55].any() and d[ages > 67].all()


@pytest.mark.parametrize("M,rng_,k", [(60.0, 6.0, 1.0), (45.0, 20.0, 0.3)])
def test_atandeath_law(M, rng_, k):
    ages = np.repeat([0.0, M - 10, M - 2, M, M + 1, M + 5, M + 40], 8000)
    d = died(population(ages, abi.ACT_ATANDEATH, atan_max_age=M, atan_range=rng_, atan_slope=k))
    s = (math.pi / 2 - 0.001) / math.atan(k * rng_)
    for a in np.unique(ages):
        m = ages == a
        p = min(1.0, max(0.0, 0.5 + s * math.atan(k * (a - M)) / math.pi))
        assert binom_ok(int(d[m].sum()), int(m.sum()), p), (a, d[m].mean(), p)


def test_death_actions_set_the_age():
    ages = np.array([10.0, 20.0, 30.0])
    for act, kw in ((abi.ACT_OLDAGEDEATH, dict(oad_max_age=100.0, oad_uncertainty=0.01)),
                    (abi.ACT_ATANDEATH, dict(atan_max_age=500.0, atan_range=1.0, atan_slope=10.0))):
        sc = population(ages, act, **kw)
        sc["agents"]["age"][:] = -1.0
        o = Oracle(sc)
        o.pre_loop()
        o.step()
        a = o.agents()
        assert np.array_equal(a["age"], ages)        # t = 0: age = 0 - birth time
