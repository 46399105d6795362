"""Model-level pins of the oracle (CPU only): invariants, closed forms and
statistical checks of the step loop semantics."""
import numpy as np
import pytest
import scipy.stats

from qhg4_amd import abi, scenarios
from qhg4_amd.grids import rect_grid, ico_grid
from oracle.oracle import Oracle, lib
from tests.helpers import assert_same_state, tut_verhulst


def _agents(ids, cells, genders, life=abi.QHG_LIFE_FERTILE, age=20.0):
    n = len(ids)
    return {"id": np.asarray(ids, dtype=np.int64), "cell": np.asarray(cells, dtype=np.int32),
            "life": np.full(n, life, dtype=np.uint8), "gender": np.asarray(genders, dtype=np.uint8),
            "birth_time": np.full(n, -age), "age": np.full(n, age), "last_birth": np.full(n, -10.0),
            "mate_id": np.full(n, -3, dtype=np.int64), "num_babies": np.zeros(n, dtype=np.int32)}


def _sc(grid, params, agents, env=None):
    return {"name": "unit", "grid": grid, "env": env or {}, "params": params, "agents": agents}


def test_schedule_independence():
    """Same state with 1 and 8 OpenMP threads (no schedule-dependent draws)."""
    states = []
    for th in (1, 8):
        lib().orc_set_threads(th)
        sc = tut_verhulst(scenarios.c1(seed=21, width=48, height=48, n_agents=3000))
        o = Oracle(sc)
        o.pre_loop()
        o.run(60)
        states.append(o.agents())
    lib().orc_set_threads(8)
    assert_same_state(states[0], states[1], "threads 1 vs 8")


def test_counts_ids_and_ages():
    sc = tut_verhulst(scenarios.c1(seed=5, width=40, height=40, n_agents=2000))
    o = Oracle(sc)
    o.pre_loop()
    max_id = 2000
    for step in range(40):
        o.step()
        a = o.agents()
        st = o.stats()
        # per-cell counts are a recount of the live agents
        assert np.array_equal(o.cell_counts(), np.bincount(a["cell"], minlength=40 * 40))
        assert st["num_alive"] == len(a["id"])
        # ids strictly increasing; new ids continue the sequence without gaps
        assert np.all(np.diff(a["id"]) > 0)
        assert st["max_id"] == max_id + st["births_last"]
        newborn = a["id"] > max_id
        assert newborn.sum() <= st["births_last"]
        max_id = st["max_id"]
        # GetOld runs last (prio 8): age = t - birth time exactly
        assert np.array_equal(a["age"], float(step) - a["birth_time"])
        # newborns: birth time t, no babies, female newborns FERTILE
        nb = a["birth_time"] == float(step)
        assert np.all(a["num_babies"][nb] == 0)
        assert np.all(a["life"][nb & (a["gender"] == 0)] == abi.QHG_LIFE_FERTILE)


def test_pairing_invariants():
    """Pairs are one female and one male of the same cell, both FERTILE at
    pairing time; each cell forms min(F, M) pairs."""
    grid = rect_grid(6, 6)
    rng = np.random.default_rng(0)
    n = 400
    cells = rng.integers(0, 36, n)
    g = rng.integers(0, 2, n)
    p = scenarios.make_params(9, [(abi.ACT_RANDOMPAIR, 1)])
    o = Oracle(_sc(grid, p, _agents(np.arange(1, n + 1), cells, g)))
    o.pre_loop()
    o.step()
    a = o.agents()
    id2i = {int(i): k for k, i in enumerate(a["id"])}
    paired = a["mate_id"] >= 0
    for k in np.flatnonzero(paired):
        m = id2i[int(a["mate_id"][k])]
        assert a["mate_id"][m] == a["id"][k]
        assert a["cell"][m] == a["cell"][k]
        assert a["gender"][m] != a["gender"][k]
    for c in range(36):
        F = int(((a["cell"] == c) & (a["gender"] == 0)).sum())
        M = int(((a["cell"] == c) & (a["gender"] == 1)).sum())
        assert int((paired & (a["cell"] == c)).sum()) == 2 * min(F, M)


def test_pairing_is_uniform_over_injections():
    """2 females, 3 males in one cell: the 6 injections F->M must be equally
    likely over seeds (a wrong k-th-unpaired selection biases this)."""
    grid = rect_grid(1, 1)
    counts = {}
    for seed in range(3000):
        p = scenarios.make_params(seed, [(abi.ACT_RANDOMPAIR, 1)])
        o = Oracle(_sc(grid, p, _agents([1, 2, 3, 4, 5], [0] * 5, [0, 0, 1, 1, 1])))
        o.pre_loop()
        o.step()
        a = o.agents()
        key = (int(a["mate_id"][0]), int(a["mate_id"][1]))
        counts[key] = counts.get(key, 0) + 1
    assert len(counts) == 6 and all(k[0] != k[1] for k in counts)
    obs = np.array(list(counts.values()))
    assert scipy.stats.chisquare(obs).pvalue > 1e-4


def test_random_move_lattice_edge_quirk():
    """Right-edge lattice cell: slot 0 (+x) is -1 and the count is 3, so
    slot = floor(u*4) in {0,1,2,3}: stay, stay (hole), +y, -x; -y is never
    reached (RandomMove.cpp:81-95 with GridFactory.cpp:979-1015)."""
    W, H = 5, 5
    grid = rect_grid(W, H)
    c = 2 * W + (W - 1)
    n = 8000
    p = scenarios.make_params(3, [(abi.ACT_RANDOMMOVE, 1)], rmove_prob=1.0)
    o = Oracle(_sc(grid, p, _agents(np.arange(1, n + 1), [c] * n, [0] * n)))
    o.pre_loop()
    o.step()
    dest = o.agents()["cell"]
    frac = {d: float(np.mean(dest == d)) for d in set(dest.tolist())}
    assert set(frac) == {c, c + W, c - 1}
    assert abs(frac[c] - 0.5) < 0.03 and abs(frac[c + W] - 0.25) < 0.03 and abs(frac[c - 1] - 0.25) < 0.03
    # interior cell: 4 neighbours + stay, 1/5 each
    ci = 2 * W + 2
    o = Oracle(_sc(grid, p, _agents(np.arange(1, n + 1), [ci] * n, [0] * n)))
    o.pre_loop()
    o.step()
    dest = o.agents()["cell"]
    obs = np.array([np.sum(dest == d) for d in (ci, ci + 1, ci + W, ci - 1, ci - W)])
    assert obs.sum() == n
    assert scipy.stats.chisquare(obs).pvalue > 1e-4


def test_weighted_move_follows_weights():
    """MultiEvaluator ADD of one identity evaluator: row = [v_c, v_n...]; a
    mover lands on slot i with probability v_i / sum (WeightedMove.cpp:64-86)."""
    grid = rect_grid(3, 3, periodic=True)
    v = np.array([1.0, 2.0, 0.0, 3.0, 4.0, 5.0, 0.0, 6.0, 0.0])
    env = {"altitude": np.full(9, 100.0), "base_npp": v}
    p = scenarios.make_params(11, [(abi.ACT_MULTIEVAL, 1), (abi.ACT_WEIGHTEDMOVE, 2)],
                              wmove_prob=1.0, me_mode=abi.ME_ADD_SIMPLE, me_num=1,
                              me_input=[abi.EV_BASE_NPP], me_cumulate=[0], me_weight=[1.0])
    n = 20000
    c = 4
    o = Oracle(_sc(grid, p, _agents(np.arange(1, n + 1), [c] * n, [0] * n), env))
    o.pre_loop()
    o.step()
    W = o.env_weights().reshape(9, 5)
    nbr = grid["neighbors"][c]
    expect_row = np.cumsum(np.concatenate([[v[c]], v[nbr]]))
    assert np.array_equal(W[c], expect_row)
    dest = o.agents()["cell"]
    targets = [c] + list(nbr)
    obs = np.array([np.sum(dest == d) for d in targets])
    probs = np.concatenate([[v[c]], v[nbr]]) / expect_row[-1]
    assert obs.sum() == n
    assert scipy.stats.chisquare(obs, probs * n).pvalue > 1e-4


def test_weighted_move_all_equal_is_uniform():
    grid = rect_grid(3, 3, periodic=True)
    env = {"altitude": np.full(9, 100.0), "base_npp": np.zeros(9)}
    p = scenarios.make_params(12, [(abi.ACT_MULTIEVAL, 1), (abi.ACT_WEIGHTEDMOVE, 2)],
                              wmove_prob=1.0, me_mode=abi.ME_ADD_SIMPLE, me_num=1,
                              me_input=[abi.EV_BASE_NPP], me_cumulate=[0], me_weight=[1.0])
    n = 10000
    o = Oracle(_sc(grid, p, _agents(np.arange(1, n + 1), [4] * n, [0] * n), env))
    o.pre_loop()
    o.step()
    dest = o.agents()["cell"]
    obs = np.array([np.sum(dest == d) for d in [4] + list(grid["neighbors"][4])])
    assert scipy.stats.chisquare(obs).pvalue > 1e-4


def test_verhulst_rates_closed_form():
    """One cell, N agents, no pairing: deaths ~ Binomial(N, p_die) with
    p_die from b = b0+(theta-b0)N/K (b<0 kills with -b) then d."""
    grid = rect_grid(1, 1)
    N, K = 3000, 1000.0
    b0, d0, th = 0.1, 0.05, 0.01
    p = scenarios.make_params(13, [(abi.ACT_VERHULST, 1)], verh_b0=b0, verh_d0=d0, verh_theta=th, verh_K=K)
    o = Oracle(_sc(grid, p, _agents(np.arange(1, N + 1), [0] * N, [1] * N)))
    o.pre_loop()
    o.step()
    b = b0 + (th - b0) * (N / K)
    d = d0 + (th - d0) * (N / K)
    p_die = (-b if b < 0 else 0) + (1 - (-b if b < 0 else 0)) * max(d, 0)
    deaths = o.stats()["deaths_last"]
    assert abs(deaths - N * p_die) < 5 * np.sqrt(N * p_die * (1 - p_die)) + 1


def test_capacity_closed_form():
    """NPPCapacity with K = Kmin + NPP*Kmax/(NPPmax-NPPmin), altitude factor
    1 below 1500 m, linear to 0 at 2500 m, sea -> 0 (NPPCapacity.cpp:152-205)."""
    grid = ico_grid(1)
    n = grid["num_cells"]
    alt = np.linspace(-500, 3000, n)
    npp = np.linspace(0.0, 2.0, n)
    p = scenarios.make_params(1, [(abi.ACT_NPPCAP, 1)], npp_min=0.0, npp_max=1.5,
                              npp_k_min=0.0, npp_k_max=75.0, npp_efficiency=1.0)
    o = Oracle(_sc(grid, p, _agents([1], [0], [0]), {"altitude": alt, "base_npp": npp}))
    o.pre_loop()
    K = o.capacities()
    af = np.clip(np.where(alt < 1500, 1.0, (2500 - alt) / 1000.0), 0, None)
    tn = npp * af
    want = np.where(alt > 0, np.where(tn > 1.5, 75.0, tn * 75.0 / 1.5), 0.0)
    assert np.allclose(K, want, rtol=1e-12, atol=1e-12)


def _pair_birth(seed, cross, rate):
    """One female (id 1) and one male (id 2) in one cell, certain birth
    (b0 = 1, K huge), no deaths; returns parent and child genomes."""
    grid = rect_grid(1, 1)
    p = scenarios.make_params(seed, [(abi.ACT_RANDOMPAIR, 1), (abi.ACT_VERHULST, 2), (abi.ACT_GENETICS, 3)],
                              verh_b0=1.0, verh_d0=0.0, verh_theta=0.0, verh_K=1e12,
                              gen_genome_size=256, gen_num_crossover=cross, gen_mutation_rate=rate)
    a = _agents([1, 2], [0, 0], [0, 1])
    rng = np.random.default_rng(seed)
    a["genome"] = rng.integers(0, 2 ** 63, size=32, dtype=np.int64).astype(np.uint64)
    o = Oracle(_sc(grid, p, a))
    o.pre_loop()
    o.step()
    out = o.agents()
    assert list(out["id"]) == [1, 2, 3]
    return a["genome"].reshape(2, 16), out["genome"][2]


def test_genetics_child_strands_from_parents():
    """Child strand 0 takes every bit from the mother's strand pair, strand 1
    from the father's (Genetics.cpp:296-320); with one crossover the strand
    switches source at most once along the 512-bit strand."""
    for seed in range(40):
        par, child = _pair_birth(seed, 1, 0.0)
        for s, parent in ((0, par[0]), (1, par[1])):
            x = child[8 * s: 8 * s + 8]
            a, b = parent[:8], parent[8:]
            assert np.all(((x ^ a) & (x ^ b)) == 0)
            # where the parent strands differ, record which one the child used
            src = []
            for w in range(8):
                diff = int(a[w] ^ b[w])
                for bit in range(64):
                    if diff >> bit & 1:
                        src.append(int(x[w] >> bit & 1) == int(a[w] >> bit & 1))
            switches = sum(1 for i in range(1, len(src)) if src[i] != src[i - 1])
            assert switches <= 1


def test_genetics_no_crossover_copies_a_strand():
    for seed in range(20):
        par, child = _pair_birth(seed, 0, 0.0)
        assert any(np.array_equal(child[:8], par[0][8 * k: 8 * k + 8]) for k in (0, 1))
        assert any(np.array_equal(child[8:], par[1][8 * k: 8 * k + 8]) for k in (0, 1))


def test_genetics_mutation_count_distribution():
    """With crossover off, the child differs from the copied strands in
    k ~ Binomial(512, rate) nucleotides (table inversion + mutateNucs)."""
    rate = 0.004
    ks = []
    for seed in range(600):
        par, child = _pair_birth(seed, 0, rate)
        best = None
        for k0 in (0, 1):
            for k1 in (0, 1):
                ref = np.concatenate([par[0][8 * k0: 8 * k0 + 8], par[1][8 * k1: 8 * k1 + 8]])
                d = ref ^ child
                m = (d | (d >> np.uint64(1))) & np.uint64(0x5555555555555555)
                nucs = sum(bin(int(w)).count("1") for w in m)
                best = nucs if best is None else min(best, nucs)
        ks.append(best)
    ks = np.array(ks)
    # mean of Binomial(512, 0.004) = 2.048; repeated positions only lower it
    assert abs(ks.mean() - 512 * rate) < 0.35
