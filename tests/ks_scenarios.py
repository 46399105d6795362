"""Scenarios of the statistical comparison with the unmodified reference.

The reference's draws come from WELL512 streams and this build's from Philox
keyed by agent id, so single runs differ; the comparison is between the
distributions over 64 seeds of the live population and the occupied-cell
count at fixed checkpoints (two-sample Kolmogorov-Smirnov at alpha = 0.01).

c1      the 128 x 128 lattice model as specified (10,000 agents, Verhulst
        b0 0.1, d0 0.05, theta 0.01, K 30): a sparse population that pairs
        rarely and dies out in about 200 steps.
growth  the same action set at the tutorial's Verhulst rates
        (QHG4/tutorial_data/xmldat/tut_Sexual.xml: b0 0.8, d0 0.001,
        theta 0.1) on a 48 x 48 lattice from 3,000 agents: a population that
        grows towards its capacity while spreading.
"""
from qhg4_amd import scenarios

SEEDS = tuple(range(1, 65))

SCENARIOS = {
    "c1": {"description": "128x128 lattice, 10000 agents, b0 0.1 d0 0.05 theta 0.01 K 30",
           "width": 128, "height": 128, "n_agents": 10000, "tut_rates": False,
           "steps": 200, "checkpoints": [5, 10, 25, 50, 75, 100, 150]},
    "growth": {"description": "48x48 lattice, 3000 agents, b0 0.8 d0 0.001 theta 0.1 K 30",
               "width": 48, "height": 48, "n_agents": 3000, "tut_rates": True,
               "steps": 150, "checkpoints": [5, 10, 25, 50, 100, 150]},
}

ALPHA = 0.01


def scenario(name, seed):
    spec = SCENARIOS[name]
    sc = scenarios.c1(seed=seed, width=spec["width"], height=spec["height"], n_agents=spec["n_agents"])
    if spec["tut_rates"]:
        p = sc["params"]
        p.verh_b0, p.verh_d0, p.verh_theta = 0.8, 0.001, 0.1
    return sc
