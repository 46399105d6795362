"""MI355X path vs the oracle, bit-exact, on the same seeded inputs."""
import numpy as np
import pytest

from qhg4_amd import abi, scenarios
from qhg4_amd.population import GpuPopulation
from oracle.oracle import Oracle
from tests.helpers import assert_same_state, tut_verhulst

pytestmark = pytest.mark.gpu


def _lockstep(sc, steps, check_every, use_run=False, capacity=0, compact_at=()):
    gpu = GpuPopulation.from_scenario(sc, capacity=capacity)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    assert_same_state(gpu.agents(), ora.agents(), "after preLoop")
    done = 0
    while done < steps:
        n = min(check_every, steps - done)
        if use_run:
            gpu.run(n)
        else:
            for s in range(n):
                t = float(done + s)
                gpu.initializeStep(t)
                gpu.doActions(0, t)
                gpu.finalizeStep()
        ora.run(n)
        done += n
        if done in compact_at:
            gpu.compactData()
        where = "%s step %d" % (sc["name"], done)
        assert_same_state(gpu.agents(), ora.agents(), where)
        assert np.array_equal(gpu.getNumAgentsArray(), ora.cell_counts()), where + ": counts"
        gs, os_ = gpu.stats(), ora.stats()
        for k in ("num_alive", "max_id", "births_last", "deaths_last", "moves_last",
                  "births_total", "deaths_total", "moves_total"):
            assert gs[k] == os_[k], "%s: stat %s %d vs %d" % (where, k, gs[k], os_[k])
        assert gs["device_error"] == 0
    return gpu, ora


def test_c1_as_specified():
    print("case: C1 128x128, 10k agents, 150 steps", flush=True)
    _lockstep(scenarios.c1(), 150, 25)


def test_c1_tutorial_rates():
    print("case: C1 lattice with tut_Sexual Verhulst values, 200 steps", flush=True)
    sc = tut_verhulst(scenarios.c1(seed=3))
    gpu, ora = _lockstep(sc, 200, 40, use_run=True)
    assert ora.stats()["num_alive"] > 10000


def test_c2_small():
    print("case: C2 model, ico level 5, 20k agents, 120 steps", flush=True)
    sc = scenarios.c2(seed=5, level=5, n_agents=20000, patch_cells=800)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    assert np.array_equal(gpu.capacities().view(np.uint64), ora.capacities().view(np.uint64))
    gpu.run(1)
    ora.run(1)
    w_g, w_o = gpu.env_weights(), ora.env_weights()
    assert np.array_equal(w_g.view(np.uint64), w_o.view(np.uint64)), "env weights differ"
    for k in range(6):
        gpu.run(20)
        ora.run(20)
        assert_same_state(gpu.agents(), ora.agents(), "C2 step %d" % (1 + 20 * (k + 1)))


def test_c3_small_genetics():
    print("case: C3 model (genetics), ico level 4, 4k agents, 80 steps", flush=True)
    sc = scenarios.c2(seed=9, level=4, n_agents=4000, patch_cells=200, genetics=True)
    sc["name"] = "C3-small"
    _lockstep(sc, 80, 20, use_run=True)


@pytest.mark.parametrize("crossovers,genome_size,rate", [(7, 256, 1e-2), (-1, 100, 5e-3), (0, 256, 2e-2),
                                                       (3, 1000, 1e-3)])
def test_genetics_variants(crossovers, genome_size, rate):
    """Several breaks per chromosome (unsorted and repeated within a block),
    free recombination, plain strand copies, genomes that do not fill their
    last word, and mutation counts above one."""
    print("case: genetics crossovers=%d genome=%d rate=%g, ico level 4, 3k agents, 40 steps"
          % (crossovers, genome_size, rate), flush=True)
    sc = scenarios.c2(seed=21 + genome_size, level=4, n_agents=3000, patch_cells=150, genetics=True)
    p = sc["params"]
    p.gen_num_crossover = crossovers
    p.gen_genome_size = genome_size
    p.gen_mutation_rate = rate
    gw = 2 * abi.genome_words(genome_size)
    rng = np.random.default_rng(genome_size)
    n = len(sc["agents"]["id"])
    sc["agents"]["genome"] = rng.integers(0, 2**63, n * gw, dtype=np.uint64)
    _lockstep(sc, 40, 10, use_run=True)


def test_compaction_keeps_state():
    print("case: forced compaction, C1 tutorial rates", flush=True)
    sc = tut_verhulst(scenarios.c1(seed=11, width=64, height=64, n_agents=3000))
    _lockstep(sc, 90, 15, compact_at=(15, 45, 60))


@pytest.mark.parametrize("genome_size", [100, 256, 300, 1000, 2048])
def test_genome_pool_compaction(genome_size):
    """Forced genome-pool compaction with genome rows shorter than a wave,
    rows that do not divide it, and rows longer than a wave (2048 loci:
    128 words)."""
    print("case: genome pool compaction, genome size %d, ico level 4, 3k agents, 36 steps" % genome_size,
          flush=True)
    sc = scenarios.c2(seed=31 + genome_size, level=4, n_agents=3000, patch_cells=150, genetics=True)
    sc["params"].gen_genome_size = genome_size
    gw = 2 * abi.genome_words(genome_size)
    rng = np.random.default_rng(genome_size)
    n = len(sc["agents"]["id"])
    sc["agents"]["genome"] = rng.integers(0, 2**63, n * gw, dtype=np.uint64)
    _lockstep(sc, 36, 6, use_run=True, compact_at=(6, 12, 24, 30))


def test_growth_beyond_capacity():
    print("case: capacity growth from a tiny initial capacity", flush=True)
    sc = tut_verhulst(scenarios.c1(seed=13, width=64, height=64, n_agents=3000))
    _lockstep(sc, 120, 30, use_run=True, capacity=1024)


def test_crowded_cells_large_pairing_paths():
    """A few cells with hundreds of fertile agents of each sex: exercises the
    in-bucket sort (> 32 per sex) and the linear k-th-unpaired selection
    (> 64 candidates) of the pairing kernel."""
    print("case: crowded cells, 6x6 lattice, 3000 agents in 6 cells", flush=True)
    sc = tut_verhulst(scenarios.c1(seed=17, width=6, height=6, n_agents=3000))
    rng = np.random.default_rng(17)
    sc["agents"]["cell"] = rng.choice([0, 7, 14, 21, 28, 35], 3000).astype(np.int32)
    sc["params"].verh_K = 2000.0
    _lockstep(sc, 30, 5)


def test_empty_and_single_agent():
    print("case: empty population and a single agent", flush=True)
    sc = scenarios.c1(seed=1, width=8, height=8, n_agents=1)
    _lockstep(sc, 10, 5)
    sc = scenarios.c1(seed=1, width=8, height=8, n_agents=0)
    gpu = GpuPopulation.from_scenario(sc)
    gpu.preLoop()
    gpu.run(5)
    assert gpu.stats()["num_alive"] == 0 and gpu.agents()["id"].size == 0


def with_action(sc, kind, prio, **kw):
    p = sc["params"]
    n = p.num_actions
    p.action_kind[n] = kind
    p.action_prio[n] = prio
    p.action_name[n].value = abi.ACTION_NAMES[kind].encode()
    p.num_actions = n + 1
    for k, v in kw.items():
        setattr(p, k, v)
    return sc


def test_c1_with_atandeath():
    """tut_Sexual-style model with ATanDeath at priority 3 (its slot in
    QHG4/tutorial_data/xmldat/tut_Sexual.xml); ages spread over the death
    ramp so both branches and deaths before pairing/moves occur."""
    sc = tut_verhulst(scenarios.c1(seed=41, width=40, height=40, n_agents=6000))
    sc["agents"]["birth_time"] = -np.linspace(0.0, 80.0, 6000)
    sc["agents"]["age"] = np.linspace(0.0, 80.0, 6000)
    with_action(sc, abi.ACT_ATANDEATH, 3, atan_max_age=45.0, atan_range=6.0, atan_slope=0.5)
    _lockstep(sc, 40, 10)


def test_c2_with_oldagedeath():
    """OoAPop's action set including OldAgeDeath (QHG4/populations/OoAPop.cpp:32-88)."""
    sc = scenarios.c2(seed=42, level=5, n_agents=20000, patch_cells=800)
    sc["agents"]["birth_time"] = -np.linspace(0.0, 70.0, 20000)
    sc["agents"]["age"] = np.linspace(0.0, 70.0, 20000)
    with_action(sc, abi.ACT_OLDAGEDEATH, 3, oad_max_age=55.0, oad_uncertainty=0.05)
    _lockstep(sc, 40, 10, use_run=True)


def test_species_xml_model():
    """A model configured from a species XML block (OoAPop with OldAgeDeath
    and the additive MultiEvaluator) runs bit-exact against the oracle."""
    import os
    from qhg4_amd import config
    xml = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "species_ooa.xml")
    sc = scenarios.c2(seed=43, level=5, n_agents=20000, patch_cells=800)
    sc["params"] = config.params_from_xml(xml, seed=43)
    sc["name"] = "species_ooa.xml"
    _lockstep(sc, 30, 10)


def test_modify_attributes_mid_run():
    """modifyAttributes between steps (RandomMove_prob, Verhulst_b0,
    ATanDeath_slope) takes effect from the next step on, as in the oracle."""
    import copy
    sc = tut_verhulst(scenarios.c1(seed=44, width=40, height=40, n_agents=5000))
    with_action(sc, abi.ACT_ATANDEATH, 10, atan_max_age=60.0, atan_range=6.0, atan_slope=1.0)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    p = copy.copy(sc["params"])
    for k, (name, field, value) in enumerate([("RandomMove_prob", "rmove_prob", 0.4),
                                              ("Verhulst_b0", "verh_b0", 0.3),
                                              ("ATanDeath_slope", "atan_slope", 0.2)]):
        gpu.run(8)
        ora.run(8)
        gpu.modifyAttributes(name, value)
        setattr(p, field, value)
        ora.set_params(p)
        assert_same_state(gpu.agents(), ora.agents(), "before change %d" % k)
    gpu.run(8)
    ora.run(8)
    assert_same_state(gpu.agents(), ora.agents(), "after the changes")
    with pytest.raises(Exception, match="environment events"):
        gpu.modifyAttributes("NPPCap_K_max", 10.0)
    gpu.close()


@pytest.mark.parametrize("genetics", [False, True, "assort"])
def test_environment_events(genetics):
    """updateEvent/flushEvents between steps: a climate event (NPP falls)
    recomputes capacities and weights; two geography events drown the agents
    of cells under water or ice (the second after the first repacked the
    buffers); stepping goes on bit-exact with the oracle.  "assort": genetics
    with AssortativePairHamming pairing."""
    print("case: environment events, ico level 4, genetics=%s" % genetics, flush=True)
    sc = scenarios.c2(seed=17, level=4, n_agents=5000, patch_cells=250, genetics=bool(genetics),
                      assort=(0.9, 1) if genetics == "assort" else None)
    gpu = GpuPopulation.from_scenario(sc)
    ora = Oracle(sc)
    gpu.preLoop()
    ora.pre_loop()
    gpu.run(6)
    ora.run(6)
    env = {k: np.array(v, copy=True) for k, v in sc["env"].items()}
    env["base_npp"] = env["base_npp"] * 0.7
    gpu.env_event({"base_npp": env["base_npp"]}, abi.EVENT_CLIMATE)
    ora.env_event({"base_npp": env["base_npp"]}, abi.EVENT_CLIMATE)
    assert np.array_equal(gpu.capacities().view(np.uint64), ora.capacities().view(np.uint64))
    assert np.array_equal(gpu.env_weights().view(np.uint64), ora.env_weights().view(np.uint64))
    rng = np.random.default_rng(8)
    for k in range(2):
        gpu.run(5)
        ora.run(5)
        occupied = np.unique(ora.agents()["cell"])
        pick = rng.permutation(occupied)
        env["altitude"][pick[:len(pick) // 5]] = -40.0
        ice = np.zeros(sc["grid"]["num_cells"], dtype=np.uint8)
        ice[pick[len(pick) // 5: len(pick) // 4]] = 1
        ev = {"altitude": env["altitude"], "ice": ice}
        n0 = ora.stats()["num_alive"]
        gpu.env_event(ev, abi.EVENT_GEO | abi.EVENT_KILL_WET)
        ora.env_event(ev, abi.EVENT_GEO | abi.EVENT_KILL_WET)
        where = "after geography event %d" % k
        assert ora.stats()["num_alive"] < n0, where
        assert_same_state(gpu.agents(), ora.agents(), where)
        assert np.array_equal(gpu.getNumAgentsArray(), ora.cell_counts()), where
        assert gpu.stats()["num_alive"] == ora.stats()["num_alive"], where
        assert np.array_equal(gpu.capacities().view(np.uint64), ora.capacities().view(np.uint64)), where
        assert np.array_equal(gpu.env_weights().view(np.uint64), ora.env_weights().view(np.uint64)), where
    for k in range(3):
        gpu.run(6)
        ora.run(6)
        assert_same_state(gpu.agents(), ora.agents(), "step block %d after the events" % k)
        gs, os_ = gpu.stats(), ora.stats()
        for key in ("num_alive", "max_id", "births_total", "deaths_total", "moves_total"):
            assert gs[key] == os_[key], key
        assert gs["device_error"] == 0
    gpu.close()


@pytest.mark.parametrize("cutoff,permute,genome_size,cells,k_max", [
    (0.9, 1, 256, 150, None),     # ~20 agents per cell, many ties, permuted
    (0.0, 0, 256, 150, None),     # every couple eligible, no permutation draws
    (0.97, 1, 1100, 150, None),   # 35 blocks: the distance spans two words per strand position
    (0.9, 1, 256, 12, 1500.0),    # 250-450 agents per cell: cells past the grouped path and LDS
])
def test_assortative_pairing(cutoff, permute, genome_size, cells, k_max):
    """OoANavGen2bitAssHamPop's AssortativePairHamming
    (QHG4/actions/AssortativePairHam.cpp:93-238, QHG4/utils/RankTable.cpp)
    on k_assort_pair, bit-exact with the oracle; such a population does not
    join a batch."""
    from qhg4_amd._lib import QhgError
    from qhg4_amd.population import GpuBatch
    sc = scenarios.c2(seed=50 + cells + genome_size, level=4, n_agents=3000, patch_cells=cells, genetics=True,
                      assort=(cutoff, permute), genome_size=genome_size)
    if k_max is not None:
        sc["params"].npp_k_max = k_max   # room for the crowded cells to live and breed
    gpu, ora = _lockstep(sc, 20, 5)
    assert gpu.stats()["births_total"] > 0
    assert (gpu.agents()["mate_id"] >= 0).any() or gpu.stats()["num_alive"] == 0
    with pytest.raises(QhgError, match="AssortativePairHamming"):
        GpuBatch().add(gpu)


@pytest.mark.parametrize("cutoff,permute,cells", [
    (1.0, 1, 150),    # at least one shared protein; ~20 agents per cell, permuted
    (0.0, 0, 150),    # every couple eligible, no permutation draws
    (2.0, 1, 40),     # crowded cells (~75 agents): cells past one wave of agents
])
def test_protein_assortative_pairing(cutoff, permute, cells):
    """OoANavGen2bitAssPop's AssortativePair
    (QHG4/actions/AssortativePair.cpp:94-271): the match of a couple is the
    number of equal MD5 digests of their proteins (ProteinBuilder,
    ProteomeComparator), then RankTable pairing; bit-exact with the oracle."""
    print("case: AssortativePair (protein matches), cutoff %s permute %d, %d patch cells" % (cutoff, permute, cells),
          flush=True)
    # seed 110: a base genome of three proteins, so that couples share 0-3
    sc = scenarios.c2(seed=110, level=4, n_agents=3000, patch_cells=cells, genetics=True,
                      assort=(cutoff, permute), protein=True)
    gpu, ora = _lockstep(sc, 20, 5)
    assert gpu.stats()["births_total"] > 0
    assert (gpu.agents()["mate_id"] >= 0).any() or gpu.stats()["num_alive"] == 0
    assert gpu.stats()["device_error"] == 0
    gpu.close()


def test_protein_budget_exceeded_fails_the_step():
    """An agent with more than 16 proteins (a genome of Met-stop codon pairs)
    exceeds the digests kept per agent: the step fails with a capacity error
    naming the limit, instead of pairing on a truncated proteome."""
    from qhg4_amd._lib import QhgError
    print("case: AssortativePair with a genome of 86 proteins", flush=True)
    sc = scenarios.c2(seed=5, level=4, n_agents=2000, patch_cells=60, genetics=True, assort=(0.0, 0), protein=True)
    # ATG TAA repeated: A=0 C=1 G=2 T=3, nucleotide i in bits 2i of its word
    codons = [(0, 3, 2), (3, 0, 0)] * 43
    nucs = [v for c in codons for v in c][:256]
    words = np.zeros(8, dtype=np.uint64)
    for i, v in enumerate(nucs):
        words[i // 32] |= np.uint64(v) << np.uint64(2 * (i % 32))
    g = sc["agents"]["genome"].reshape(len(sc["agents"]["id"]), -1)
    g[:, :8] = words
    g[:, 8:] = words
    sc["agents"]["genome"] = np.ascontiguousarray(g.ravel())
    a = sc["agents"]
    a["life"][:] = abi.QHG_LIFE_FERTILE
    a["age"][:] = 20.0
    gpu = GpuPopulation.from_scenario(sc)
    gpu.preLoop()
    n0 = gpu.getNumAgentsEffective()
    gpu.run(2)
    st = gpu.stats()
    assert st["device_error"] & 16, st["device_error"]   # the step froze the population
    assert st["num_alive"] == n0 and st["births_total"] == 0
    # the host reports it once its status catches up (the status ring holds 8 steps)
    with pytest.raises(QhgError, match="more than 16 proteins"):
        gpu.run(20)
    gpu.close()
