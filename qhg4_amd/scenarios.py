"""Scenario builders: grid + environment + parameters + initial agents.

C1  plain model on a 128x128 non-periodic lattice (tut_SexualPop action set,
    QHG4/populations/tut_SexualPop.cpp:24-43, priorities of
    QHG4/tutorial_data/xmldat/tut_Sexual.xml without ATanDeath).
C2  OoA-style model on a 40,962-cell icosahedral grid (OoAPop action set,
    QHG4/populations/OoAPop.cpp:32-88, priorities of
    QHG4/tutorial_data/xmldat/tut_EnvironCapAlt.xml, MultiEvaluator in
    multiplicative mode, no OldAgeDeath).
C3  C2 plus 2-bit diploid genetics (OoANavGen2bitPop,
    QHG4/populations/OoANavGen2bitPop.cpp:75-166, 226-241).
C4  an ensemble of C2-style replicas of uneven grid and population size with
    per-replica parameters (c4_replica).
C5  a stream of C2-style scenario requests (c5_requests / c5_scenario).

All inputs are synthetic and seeded; the same dict feeds the MI355X path and
the oracle.
"""
import numpy as np

from . import abi
from .grids import rect_grid, ico_grid, value_noise, bfs_patch, bfs_levels_patch

ALT_SIGMOID_THRESHOLD = 2500.0
ALT_SIGMOID_WIDTH = 400.0


def make_params(seed, actions, **kw):
    """actions: list of (kind, prio[, name]); kw: qhg_params fields."""
    p = abi.Params()
    p.seed = int(seed) & 0xFFFFFFFFFFFFFFFF
    p.num_actions = len(actions)
    for i, a in enumerate(actions):
        kind, prio = a[0], a[1]
        name = a[2] if len(a) > 2 else abi.ACTION_NAMES[kind]
        p.action_kind[i] = kind
        p.action_prio[i] = prio
        p.action_name[i].value = name.encode()
    p.npp_efficiency = 1.0
    for k, v in kw.items():
        if k in ("me_input", "me_cumulate", "me_weight", "me_pl_n"):
            arr = getattr(p, k)
            for i, x in enumerate(v):
                arr[i] = x
        elif k == "me_pl":
            for e, pts in enumerate(v):
                p.me_pl_n[e] = len(pts)
                for j, (x, y) in enumerate(pts):
                    p.me_pl_x[e][j] = x
                    p.me_pl_v[e][j] = y
        else:
            setattr(p, k, v)
    return p


def sigmoid_polyline(lo=-4000.0, hi=6000.0, npts=41):
    """Altitude preference 1/(1+exp((a-2500)/400)) sampled as PolyLine points."""
    xs = np.linspace(lo, hi, npts)
    return [(float(x), float(1.0 / (1.0 + np.exp((x - ALT_SIGMOID_THRESHOLD) / ALT_SIGMOID_WIDTH))))
            for x in xs]


def _agents(rng, cells, n, life, last_birth, first_id=1):
    age = rng.uniform(0.0, 40.0, n)
    return {
        "id": np.arange(first_id, first_id + n, dtype=np.int64),
        "cell": cells.astype(np.int32),
        "life": np.full(n, life, dtype=np.uint8),
        "gender": (rng.random(n) < 0.5).astype(np.uint8),
        "birth_time": -age,
        "age": age,
        "last_birth": np.full(n, last_birth, dtype=np.float64),
        "mate_id": np.full(n, -3, dtype=np.int64),
        "num_babies": np.zeros(n, dtype=np.int32),
    }


def c1(seed=42, width=128, height=128, n_agents=10000):
    """Plain model, small lattice.  No Fertility limits are named for C1, so
    Fertility runs with min age 0 and no upper bound: every agent older than
    0 can pair, which is the 'uniform random female-male pairing'."""
    grid = rect_grid(width, height, periodic=False)
    rng = np.random.default_rng(seed)
    cells = rng.integers(0, grid["num_cells"], n_agents)
    agents = _agents(rng, cells, n_agents, abi.QHG_LIFE_FERTILE, 0.0)
    params = make_params(
        seed,
        [(abi.ACT_FERTILITY, 2), (abi.ACT_RANDOMPAIR, 3), (abi.ACT_VERHULST, 6),
         (abi.ACT_RANDOMMOVE, 7), (abi.ACT_GETOLD, 8)],
        fert_min_age=0.0, fert_max_age=1e30, fert_interbirth=-1.0,
        verh_b0=0.1, verh_d0=0.05, verh_theta=0.01, verh_K=30.0,
        rmove_prob=0.1)
    return {"name": "C1", "grid": grid, "env": {}, "params": params, "agents": agents,
            "steps": 1000}


def ico_env(grid, seed):
    xyz = grid["xyz"]
    alt = -4000.0 + 10000.0 * value_noise(xyz, seed * 2 + 1)
    npp = np.clip(value_noise(xyz, seed * 2 + 2) * 1.2, 0.0, 1.5)
    return {"altitude": alt, "base_npp": npp, "longitude": grid["lon"], "latitude": grid["lat"]}


def ooa_params(seed, genetics=False, wmove_prob=0.2, b0=0.8, d0=0.001, theta=0.1, k_scale=50.0, assort=None,
               genome_size=256, protein=False, navigate=None):
    """VerhulstVarK values are those of QHG4/tutorial_data/xmldat/tut_EnvironCapAlt.xml
    (C2 names none; with C1's b0=0.1, d0=0.05 the paired-female birth rate
    cannot outpace deaths and every run dies out).  assort=(cutoff, permute)
    pairs with AssortativePairHamming instead of RandomPair, as
    OoANavGen2bitAssHamPop does (needs genetics); with protein, with
    AssortativePair (protein matches), as OoANavGen2bitAssPop does."""
    pair = abi.ACT_RANDOMPAIR if assort is None else (abi.ACT_ASSORTPROT if protein else abi.ACT_ASSORTPAIR)
    # the genetics classes (OoANavGen2bitPop, OoANavGen2bitAssHamPop) read
    # their NPP evaluator through a PolyLine (NPPPref; a SingleEvaluator with
    # a PolyLine name needs at least two points, PolyLine.cpp:92-127): a ramp
    # through (0, 0) and (K_max, K_max) passes the capacity on, as a float;
    # OoAPop's NPP evaluator reads the capacity directly
    k_max = 1.5 * k_scale
    npp_pl = [(0.0, 0.0), (k_max, k_max)] if genetics else []
    actions = [(abi.ACT_NPPCAP, 1), (abi.ACT_GETOLD, 2), (abi.ACT_WEIGHTEDMOVE, 4),
               (abi.ACT_MULTIEVAL, 5, "MultiEvaluator[Alt+NPP]"), (abi.ACT_FERTILITY, 6),
               (pair, 7), (abi.ACT_VERHULSTVARK, 8)]
    kw = {}
    if genetics:
        actions.append((abi.ACT_GENETICS, 9))
        kw = dict(gen_genome_size=genome_size, gen_num_crossover=1, gen_mutation_rate=1e-3)
    if assort is not None:
        kw.update(assort_cutoff=float(assort[0]), assort_permute=int(assort[1]))
    if navigate is not None:
        prio, decay, dist0, prob0, bridge = navigate
        actions.append((abi.ACT_NAVIGATE, prio))
        kw.update(nav_decay=decay, nav_dist0=dist0, nav_prob0=prob0, nav_min_dens=0.0, nav_bridge_prob=bridge)
    return make_params(
        seed, actions,
        fert_min_age=15.0, fert_max_age=45.0, fert_interbirth=3.0,
        verh_b0=b0, verh_d0=d0, verh_theta=theta,
        wmove_prob=wmove_prob,
        me_mode=abi.ME_MUL_SIMPLE, me_num=2,
        me_input=[abi.EV_ALTITUDE, abi.EV_CAPACITY], me_cumulate=[0, 0], me_weight=[1.0, 1.0],
        me_pl=[sigmoid_polyline(), npp_pl],
        npp_min=0.0, npp_max=1.5, npp_k_min=0.0, npp_k_max=k_max, npp_efficiency=1.0,
        **kw)


def c2(seed=42, level=6, n_agents=100000, patch_cells=4000, genetics=False, assort=None, genome_size=256,
       protein=False):
    """100k agents in a BFS patch of land cells.  The patch is 4,000 cells
    (~25 agents per cell, about half of K): a 200-cell patch puts 500 agents
    in cells with K <= 75, where b = b0 + (theta-b0)*N/K < -1 and the
    Verhulst rule kills every agent in step 0."""
    grid = ico_grid(level)
    env = ico_env(grid, seed)
    alt, npp = env["altitude"], env["base_npp"]
    land = alt > 0
    good = land & (alt < 1500)
    start = int(np.argmax(np.where(good, npp, -1.0)))
    patch = bfs_patch(grid, start, patch_cells, land)
    rng = np.random.default_rng(seed + 7)
    cells = patch[rng.integers(0, len(patch), n_agents)]
    agents = _agents(rng, cells, n_agents, abi.QHG_LIFE_ALIVE, -1.0)
    params = ooa_params(seed, genetics=genetics, assort=assort, genome_size=genome_size, protein=protein)
    out = {"name": "C2", "grid": grid, "env": env, "params": params, "agents": agents,
           "steps": 2000}
    if genetics:
        out["agents"]["genome"] = initial_genomes(rng, n_agents, genome_size)
    return out


EARTH_RADIUS_KM = 6371.3


def c2nav(seed=42, level=5, n_agents=20000, patch_cells=800, nav_first=False, prob0=0.1, bridge_prob=0.3,
          genetics=True):
    """C2-style model with Navigate (OoANavGen2bitPop schedules it,
    QHG4/populations/OoANavGen2bitPop.cpp:77-93), on a coast: the agents
    start in a patch of land around the coastal cell with the best NPP.
    Ports are the patch's coastal cells plus random coastal cells elsewhere;
    each port gets 0-6 destinations among the 40 coastal cells nearest to it,
    at great-circle distances in km.  Bridges join ports to patch cells, to
    water cells (never current while the water stays) and one cell to
    itself.  3 % of the cells are ice.  With nav_first, Navigate shares
    WeightedMove's priority and runs before it (name order), so a later
    WeightedMove registration can override a sea-way jump; otherwise it runs
    after WeightedMove and only agents that did not move navigate."""
    grid = ico_grid(level)
    env = dict(ico_env(grid, seed))
    alt = env["altitude"]
    nbr = grid["neighbors"]
    land = alt > 0
    wet_nb = np.array([any(nb >= 0 and alt[nb] <= 0 for nb in row) for row in nbr])
    coastal = np.flatnonzero(land & wet_nb)
    score = np.where(np.isin(np.arange(grid["num_cells"]), coastal) & (alt < 1500), env["base_npp"], -1.0)
    start = int(np.argmax(score))
    patch = bfs_patch(grid, start, patch_cells, land)
    rng = np.random.default_rng(seed + 23)
    in_patch = np.zeros(grid["num_cells"], dtype=bool)
    in_patch[patch] = True
    ports = sorted(set(coastal[in_patch[coastal]].tolist()) |
                   set(rng.choice(coastal, size=min(len(coastal), 40), replace=False).tolist()))
    xyz = grid["xyz"]
    xyz_c = xyz[coastal] / np.linalg.norm(xyz[coastal], axis=1)[:, None]
    dests = {}
    for c in ports:
        u = xyz[c] / np.linalg.norm(xyz[c])
        ang = np.arccos(np.clip(xyz_c @ u, -1.0, 1.0))
        near = coastal[np.argsort(ang)[1:41]]
        k = int(rng.integers(0, 7))
        pick = rng.choice(near, size=k, replace=False)
        dests[int(c)] = [(int(d), float(EARTH_RADIUS_KM * np.arccos(np.clip(
            (xyz[d] / np.linalg.norm(xyz[d])) @ u, -1.0, 1.0)))) for d in pick]
    water = np.flatnonzero(~land)
    bridges = [(int(rng.choice(ports)), int(rng.choice(patch))) for _ in range(20)]
    bridges += [(int(rng.choice(patch)), int(rng.choice(water))) for _ in range(4)]
    bridges.append((int(patch[0]), int(patch[0])))
    ice = np.zeros(grid["num_cells"], dtype=np.uint8)
    ice[rng.random(grid["num_cells"]) < 0.03] = 1
    env["ice"] = ice
    cells = patch[rng.integers(0, len(patch), n_agents)]
    agents = _agents(rng, cells, n_agents, abi.QHG_LIFE_ALIVE, -1.0)
    nav_prio = 4 if nav_first else 5
    params = ooa_params(seed, genetics=genetics, navigate=(nav_prio, -0.001, 150.0, prob0, bridge_prob))
    out = {"name": "C2-Navigate", "grid": grid, "env": env, "params": params, "agents": agents,
           "nav": abi.nav_arrays(dests, bridges, sample_dist=1.0), "steps": 200}
    if genetics:
        out["agents"]["genome"] = initial_genomes(rng, n_agents, 256)
    return out


def initial_genomes(rng, n, genome_size, muts=8):
    """A shared random base genome with ``muts`` random nucleotide changes
    per agent (stand-in for GenomeCreator; input data only)."""
    w = abi.genome_words(genome_size)
    base = rng.integers(0, 2 ** 63, size=2 * w, dtype=np.int64).astype(np.uint64)
    g = np.tile(base, (n, 1))
    pos = rng.integers(0, 2 * genome_size, size=(n, muts))
    val = rng.integers(1, 4, size=(n, muts)).astype(np.uint64)
    rows = np.repeat(np.arange(n), muts)
    word = (pos // 32).ravel()
    shift = (2 * (pos % 32)).ravel().astype(np.uint64)
    np.bitwise_xor.at(g, (rows, word), val.ravel() << shift)
    return np.ascontiguousarray(g.ravel())


def c3(seed=42, level=8, n_agents=2_000_000):
    """Scale run with genetics: agents spread uniformly over land cells below
    1500 m."""
    grid = ico_grid(level)
    env = ico_env(grid, seed)
    alt = env["altitude"]
    hab = np.flatnonzero((alt > 0) & (alt < 1500) & (env["base_npp"] > 0.05)).astype(np.int32)
    rng = np.random.default_rng(seed + 11)
    cells = hab[rng.integers(0, len(hab), n_agents)]
    agents = _agents(rng, cells, n_agents, abi.QHG_LIFE_ALIVE, -1.0)
    agents["genome"] = initial_genomes(rng, n_agents, 256)
    params = ooa_params(seed, genetics=True)
    return {"name": "C3", "grid": grid, "env": env, "params": params, "agents": agents,
            "steps": 5000}


# -- C4: ensemble of C2-style replicas ----------------------------------------
C4_LEVELS = (4, 5, 6, 7)          # 2,562 / 10,242 / 40,962 / 163,842 cells
C4_N0 = (10_000, 2_000_000)       # initial agents, log-uniform
C4_DENSITY = 25.0                 # initial agents per land cell (as C2)
_c4_worlds = {}


def c4_world(level, base_seed):
    """Grid and environment of one C4 grid size, shared by every replica of
    that size (the fields are seeded by the ensemble's base seed)."""
    key = (level, base_seed)
    if key not in _c4_worlds:
        grid = ico_grid(level)
        env = ico_env(grid, base_seed + level)
        alt = env["altitude"]
        land = alt > 0
        good = np.flatnonzero(land & (alt < 1500) & (env["base_npp"] > 0.05)).astype(np.int32)
        _c4_worlds[key] = (grid, env, land, good)
    return _c4_worlds[key]


def c4_draw(r, base_seed=1000):
    """The drawn sizes and parameters of replica r (seed base_seed + r):
    N0 log-uniform in [10k, 2M]; the grid uniformly among the C4 sizes whose
    land holds N0 at 25 agents per cell (so no replica starts far above its
    carrying capacity and dies out in its first step); move probability
    U(0.05, 0.5), b0 U(0.05, 0.3), d0 U(0.01, 0.1), K scale U(20, 80)."""
    rng = np.random.default_rng(base_seed + r)
    n0 = int(round(np.exp(rng.uniform(np.log(C4_N0[0]), np.log(C4_N0[1])))))
    fits = [lv for lv in C4_LEVELS if 0.8 * (10 * 4 ** lv + 2) * C4_DENSITY >= n0] or [C4_LEVELS[-1]]
    level = fits[int(rng.integers(0, len(fits)))]
    return {
        "n0": n0, "level": level,
        "wmove_prob": float(rng.uniform(0.05, 0.5)), "b0": float(rng.uniform(0.05, 0.3)),
        "d0": float(rng.uniform(0.01, 0.1)), "k_scale": float(rng.uniform(20.0, 80.0)),
        "start": float(rng.random()), "seed": base_seed + r,
    }


def c4_replica(r, base_seed=1000, steps=1000):
    """Replica r of the C4 ensemble: the C2 model (same actions, priorities
    and evaluators) with the drawn sizes and parameters; N0 agents in a BFS
    patch of land cells at 25 agents per cell around a random habitable cell."""
    sc = scenario_from_draw(c4_draw(r, base_seed), base_seed)
    sc["name"] = "C4[%d]" % r
    sc["steps"] = steps
    return sc


# -- C5: served scenario stream ------------------------------------------------
C5_LEVELS = (4, 5, 6, 7, 8)       # 2,562 ... 655,362 cells
C5_N0 = (10_000, 5_000_000)
C5_HORIZON = (200, 5000)
C5_RATE = 20.0                    # requests per second (Poisson)


def c5_requests(n, base_seed=2000, rate=C5_RATE):
    """n requests: arrival times of a Poisson process and each request's
    drawn sizes, horizon and parameters (seed base_seed + i)."""
    rng = np.random.default_rng(base_seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, n))
    out = []
    for i in range(n):
        r = np.random.default_rng(base_seed + 1 + i)
        n0 = int(round(np.exp(r.uniform(np.log(C5_N0[0]), np.log(C5_N0[1])))))
        fits = [lv for lv in C5_LEVELS if 0.8 * (10 * 4 ** lv + 2) * C4_DENSITY >= n0] or [C5_LEVELS[-1]]
        out.append({
            "id": i, "arrival": float(arrivals[i]), "n0": n0, "level": fits[int(r.integers(0, len(fits)))],
            "horizon": int(r.integers(C5_HORIZON[0], C5_HORIZON[1] + 1)),
            "wmove_prob": float(r.uniform(0.05, 0.5)), "b0": float(r.uniform(0.05, 0.3)),
            "d0": float(r.uniform(0.01, 0.1)), "k_scale": float(r.uniform(20.0, 80.0)),
            "start": float(r.random()), "seed": base_seed + 1 + i,
        })
    return out


def scenario_from_draw(d, world_seed):
    """A C2-model scenario from a drawn request/replica: N0 agents at 25 per
    land cell in a BFS patch around a habitable cell of the drawn grid size."""
    grid, env, land, good = c4_world(d["level"], world_seed)
    start = int(good[int(d["start"] * len(good))])
    patch = bfs_levels_patch(grid, start, int(np.ceil(d["n0"] / C4_DENSITY)), land)
    rng = np.random.default_rng(d["seed"] + 7)
    cells = patch[rng.integers(0, len(patch), d["n0"])]
    agents = _agents(rng, cells, d["n0"], abi.QHG_LIFE_ALIVE, -1.0)
    # listed by cell (ids ascending within a cell): the same agents, in the
    # order the device keeps them, so the upload has nothing to sort
    order = np.argsort(agents["cell"], kind="stable")
    agents = {k: v[order] for k, v in agents.items()}
    params = ooa_params(d["seed"], wmove_prob=d["wmove_prob"], b0=d["b0"], d0=d["d0"], k_scale=d["k_scale"])
    return {"grid": grid, "env": env, "params": params, "agents": agents, "draw": d}


def c5_scenario(req, world_seed=2000):
    sc = scenario_from_draw(req, world_seed)
    sc["name"] = "C5[%d]" % req["id"]
    sc["steps"] = req["horizon"]
    return sc


def grid_desc(grid):
    """GridDesc struct pointing into the grid dict's arrays (keep the dict alive)."""
    g = abi.GridDesc()
    g.num_cells = int(grid["num_cells"])
    g.connectivity = int(grid["connectivity"])
    g.neighbors = abi.ptr(grid["neighbors"], abi.C.c_int32)
    g.num_neighbors = abi.ptr(grid["num_neighbors"], abi.C.c_uint8)
    g.cell_ids = abi.ptr(grid.get("cell_ids"), abi.C.c_int64)
    return g


def env_desc(env):
    e = abi.EnvDesc()
    for k in ("altitude", "longitude", "latitude", "water", "base_npp", "temperature", "rainfall"):
        arr = env.get(k)
        if arr is not None:
            arr = env[k] = np.ascontiguousarray(arr, dtype=np.float64)
        setattr(e, k, abi.ptr(arr, abi.C.c_double))
    for k in ("ice", "coastal"):
        setattr(e, k, abi.ptr(env.get(k), abi.C.c_uint8))
    return e
