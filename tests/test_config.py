"""Species configuration from the reference's XML (CPU: no GPU calls).

The tutorial species blocks (tests/golden/species_tutorial.json, written by
tests/golden/make_species_fixtures.py from QHG4/tutorial_data/xmldat) must
yield parameters holding exactly their values; the rules of
Action::checkAttributes / Prioritizer (missing or unknown names, priorities
for absent actions) must reject what the reference rejects."""
import copy
import json
import os

import numpy as np
import pytest

from oracle.oracle import Oracle
from qhg4_amd import abi, config, scenarios
from qhg4_amd._lib import QhgError
from qhg4_amd.grids import ico_grid

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "species_tutorial.json")))
REF_XML = "/root/reference/QHG4/tutorial_data/xmldat"


def spec(name):
    s = copy.deepcopy(FIX["species"][name])
    s["modules"] = [tuple(m) for m in s["modules"]]
    s["priorities"] = [tuple(p) for p in s["priorities"]]
    return s


def actions(p):
    return {p.action_name[i].value.decode(): (p.action_kind[i], p.action_prio[i]) for i in range(p.num_actions)}


def test_tut_sexual_species():
    p = config.params_from_species(spec("tut_Sexual"), seed=7)
    assert p.seed == 7
    assert actions(p) == {"GetOld": (abi.ACT_GETOLD, 8), "ATanDeath": (abi.ACT_ATANDEATH, 10),
                          "RandomMove": (abi.ACT_RANDOMMOVE, 7), "Fertility": (abi.ACT_FERTILITY, 2),
                          "RandomPair": (abi.ACT_RANDOMPAIR, 3), "Verhulst": (abi.ACT_VERHULST, 6)}
    for name, params in spec("tut_Sexual")["modules"]:
        for k, v in params.items():
            assert config.get_attribute(p, k) == float(v), (name, k)
    assert (p.atan_max_age, p.atan_range, p.atan_slope) == (60.0, 6.0, 1.0)
    assert (p.verh_b0, p.verh_d0, p.verh_theta, p.verh_K) == (0.8, 0.001, 0.1, 30.0)


def test_tut_environcapalt_species():
    p = config.params_from_species(spec("tut_EnvironCapAlt"), seed=1)
    a = actions(p)
    assert a["MultiEvaluator[NPP+Alt]"] == (abi.ACT_MULTIEVAL, 5)
    assert a["NPPCapacity"] == (abi.ACT_NPPCAP, 1) and a["VerhulstVarK"] == (abi.ACT_VERHULSTVARK, 8)
    # MultiEvaluator wiring of tut_EnvironCapAltPop: Alt (PolyLine AltPref) + NPP (capacity, direct)
    assert p.me_mode == abi.ME_ADD_SIMPLE and p.me_num == 2
    assert list(p.me_input[:2]) == [abi.EV_ALTITUDE, abi.EV_CAPACITY]
    assert list(p.me_weight[:2]) == [0.2, 0.8]
    assert p.me_pl_n[0] == 5 and p.me_pl_n[1] == 0
    assert list(p.me_pl_x[0][:5]) == [-0.1, 0.1, 1500.0, 2000.0, 3000.0]
    assert list(p.me_pl_v[0][:5]) == [0.0, 0.01, 1.0, 1.0, -9999.0]
    assert (p.npp_k_max, p.npp_max) == (38.4716796875, 1.0576171875)
    assert (p.npp_water_factor, p.npp_coastal_factor, p.npp_efficiency) == (0.5349609375, 0.4, 1.0)
    assert (p.fert_min_age, p.fert_max_age, p.fert_interbirth) == (15.0, 50.0, 2.0)
    assert p.wmove_prob == 0.07


@pytest.mark.skipif(not os.path.isdir(REF_XML), reason="reference tutorial data not present")
def test_reader_reproduces_the_fixture_from_the_reference_files():
    for name in ("tut_Sexual", "tut_EnvironCapAlt"):
        got = config.read_species_xml(os.path.join(REF_XML, name + ".xml"))
        assert json.loads(json.dumps(got)) == FIX["species"][name]


def test_own_ooa_species_runs_on_the_oracle():
    """An OoAPop block (with OldAgeDeath, as OoAPop.cpp:32-88 wires it)
    gives a runnable model."""
    p = config.params_from_xml(os.path.join(HERE, "golden", "species_ooa.xml"), seed=3)
    a = actions(p)
    assert a["OldAgeDeath"] == (abi.ACT_OLDAGEDEATH, 3) and a["MultiEvaluator[Alt+NPP]"][1] == 5
    assert list(p.me_input[:2]) == [abi.EV_ALTITUDE, abi.EV_CAPACITY] and p.me_pl_n[0] == 5
    sc = scenarios.c2(seed=3, level=4, n_agents=3000, patch_cells=150)
    sc["params"] = p
    o = Oracle(sc)
    o.pre_loop()
    o.run(10)
    assert o.stats()["step"] == 10 and o.stats()["num_alive"] > 0


def edit(s, module, key=None, value=None, drop=None):
    s = copy.deepcopy(s)
    for i, (n, prm) in enumerate(s["modules"]):
        if n == module:
            prm = dict(prm)
            if drop:
                prm.pop(drop)
            if key:
                prm[key] = value
            s["modules"][i] = (n, prm)
    return s


@pytest.mark.parametrize("mutate,msg", [
    (lambda s: edit(s, "Verhulst", drop="Verhulst_K"), "missing parameter Verhulst_K"),
    (lambda s: edit(s, "Verhulst", key="Verhulst_Q", value="1"), "unknown parameter Verhulst_Q"),
    (lambda s: edit(s, "Fertility", key="Fertility_min_age", value="fourteen"), "bad value"),
    (lambda s: dict(s, modules=[m for m in s["modules"] if m[0] != "ATanDeath"]), "no module for ATanDeath"),
    (lambda s: dict(s, priorities=s["priorities"] + [("WeightedMove", 4)]), "does not have"),
    (lambda s: dict(s, priorities=s["priorities"] + [("GetOld", 1)]), "twice"),
    (lambda s: dict(s, **{"class": "NoSuchPop"}), "unknown population class"),
])
def test_rejects_what_the_reference_rejects(mutate, msg):
    with pytest.raises(QhgError, match=msg):
        config.params_from_species(mutate(spec("tut_Sexual")), seed=1)


def test_multievaluator_errors():
    s = spec("tut_EnvironCapAlt")
    with pytest.raises(QhgError, match="Multi_weight_npp"):
        config.params_from_species(edit(s, "MultiEvaluator[NPP+Alt]", drop="Multi_weight_npp"), 1)
    with pytest.raises(QhgError, match="PolyLine AltPref"):
        config.params_from_species(edit(s, "SingleEvaluator[Alt]", key="AltPref", value="1 2 3"), 1)


def test_unscheduled_actions_are_not_run():
    s = spec("tut_Sexual")
    s["priorities"] = [p for p in s["priorities"] if p[0] != "ATanDeath"]
    p = config.params_from_species(s, 1)
    assert "ATanDeath" not in actions(p) and p.atan_max_age == 60.0


NAV_MODULE = """<module name="Navigate">
    <param name="Navigate_decay" value="-0.001"/><param name="Navigate_dist0" value="150.0"/>
    <param name="Navigate_prob0" value="0.1"/><param name="Navigate_min_dens" value="0.0"/>
    <param name="Navigate_bridge_prob" value="0.25"/>
  </module>
  """


def test_navigate_species():
    """OoANavGen2bitPop holds Navigate (QHG4/populations/OoANavGen2bitPop.cpp
    :77-93): its module must be there with all five attributes, as
    Action::checkAttributes demands them (QHG4/actions/Action.cpp:54-75,
    Navigate.cpp:18-24); it runs only with a priority."""
    xml = open(os.path.join(HERE, "golden", "species_ooa.xml")).read()
    xml = xml.replace('name="OoAPop"', 'name="OoANavGen2bitPop"')
    xml = xml.replace("<priorities>", '''<module name="Genetics">
    <param name="Genetics_genome_size" value="256"/><param name="Genetics_num_crossover" value="1"/>
    <param name="Genetics_mutation_rate" value="0.001"/><param name="Genetics_initial_muts" value="random"/>
    <param name="Genetics_create_new_genome" value="1"/><param name="Genetics_bits_per_nuc" value="2"/>
  </module>
  ''' + NAV_MODULE + '''<priorities><prio name="Genetics" value="9"/>''')
    xml = xml.replace('id="NPP">', 'id="NPP"><param name="NPPPref" value="0 0 1.5 1"/>')
    p = config.params_from_xml(xml, seed=2)
    assert actions(p)["Genetics"] == (abi.ACT_GENETICS, 9) and p.gen_genome_size == 256
    assert p.me_pl_n[1] == 2
    assert "Navigate" not in actions(p)
    assert (p.nav_decay, p.nav_dist0, p.nav_prob0, p.nav_min_dens, p.nav_bridge_prob) == (-0.001, 150.0, 0.1, 0.0, 0.25)
    q = config.params_from_xml(xml.replace("<priorities>", '<priorities><prio name="Navigate" value="4"/>'), 2)
    assert actions(q)["Navigate"] == (abi.ACT_NAVIGATE, 4)
    assert config.get_attribute(q, "Navigate_prob0") == 0.1
    with pytest.raises(QhgError, match="no module for Navigate"):
        config.params_from_xml(xml.replace(NAV_MODULE, ""), 2)
    with pytest.raises(QhgError, match="missing parameter Navigate_bridge_prob"):
        config.params_from_xml(xml.replace('<param name="Navigate_bridge_prob" value="0.25"/>', ""), 2)
    with pytest.raises(QhgError, match="2-bit"):
        config.params_from_xml(xml.replace('"Genetics_bits_per_nuc" value="2"', '"Genetics_bits_per_nuc" value="1"'),
                               2)


@pytest.mark.parametrize("cls,act,kind,cutoff", [
    ("OoANavGen2bitAssHamPop", "AssortativePairHamming", abi.ACT_ASSORTPAIR, 0.95),
    ("OoANavGen2bitAssPop", "AssortativePair", abi.ACT_ASSORTPROT, 1.0),
])
def test_assortative_species(cls, act, kind, cutoff):
    """OoANavGen2bitAssHamPop (QHG4/populations/OoANavGen2bitAssHamPop.cpp:31-93)
    pairs with AssortativePairHamming and OoANavGen2bitAssPop
    (QHG4/populations/OoANavGen2bitAssPop.cpp:33-95) with AssortativePair;
    each module holds exactly its two attributes
    (QHG4/actions/AssortativePairHam.h:11-13, AssortativePair.h:11-13)."""
    xml = open(os.path.join(HERE, "golden", "species_ooa.xml")).read()
    xml = xml.replace('name="OoAPop"', 'name="%s"' % cls)
    xml = xml.replace('<prio name="RandomPair" value="7"/>', '<prio name="%s" value="7"/>' % act)
    xml = xml.replace("<priorities>", '''<module name="Genetics">
    <param name="Genetics_genome_size" value="256"/><param name="Genetics_num_crossover" value="1"/>
    <param name="Genetics_mutation_rate" value="0.001"/><param name="Genetics_initial_muts" value="random"/>
    <param name="Genetics_create_new_genome" value="1"/><param name="Genetics_bits_per_nuc" value="2"/>
  </module>
  <module name="%s">
    <param name="%s_cutoff" value="%s"/><param name="%s_permute" value="1"/>
  </module>
  ''' % (act, act, cutoff, act) + NAV_MODULE + '<priorities><prio name="Genetics" value="9"/>')
    xml = xml.replace('id="NPP">', 'id="NPP"><param name="NPPPref" value="0 0 1.5 1"/>')
    p = config.params_from_xml(xml, seed=2)
    a = actions(p)
    assert a[act] == (kind, 7) and "RandomPair" not in a
    assert (p.assort_cutoff, p.assort_permute) == (cutoff, 1)
    assert config.get_attribute(p, act + "_cutoff") == cutoff
    with pytest.raises(QhgError, match="missing parameter %s_permute" % act):
        config.params_from_xml(xml.replace('<param name="%s_permute" value="1"/>' % act, ""), 2)
    with pytest.raises(QhgError, match="does not have"):
        config.params_from_xml(xml.replace("<priorities>", '<priorities><prio name="RandomPair" value="3"/>'), 2)
    # the model runs on the oracle
    sc = scenarios.c2(seed=2, level=4, n_agents=3000, patch_cells=150, genetics=True)
    sc["params"] = p
    o = Oracle(sc)
    o.pre_loop()
    o.run(8)
    assert o.stats()["births_total"] > 0


def test_attribute_get_set_round_trip():
    p = config.params_from_species(spec("tut_Sexual"), 1)
    from qhg4_amd._lib import lib
    assert lib().qhg_params_set_attribute(p, b"Verhulst_b0", 0.55) == 0
    assert config.get_attribute(p, "Verhulst_b0") == 0.55 and p.verh_b0 == 0.55
    assert lib().qhg_params_set_attribute(p, b"Genetics_genome_size", 3.0) != 0   # an int attribute
    with pytest.raises(QhgError):
        config.get_attribute(p, "NoSuch_attr")


def test_agent_dat_rows_land_in_the_nearest_cell():
    rows = FIX["agents"]["tut_Sexual"]
    grid = ico_grid(4)
    a = config.read_agent_dat("\n".join(rows) + "\n", grid)
    n = len(rows) - 1
    assert len(a["id"]) == n and list(a["id"]) == list(range(1, n + 1))
    assert set(a["gender"].tolist()) <= {0, 1} and (a["life"] == 1).all()
    assert np.array_equal(a["birth_time"], np.full(n, -10.0)) and np.array_equal(a["age"], np.full(n, 10.0))
    # the row position is 8E 47N: the chosen cell is the nearest by great-circle distance
    lon, lat = np.radians(grid["lon"]), np.radians(grid["lat"])
    d = np.arccos(np.clip(np.sin(np.radians(47)) * np.sin(lat) +
                          np.cos(np.radians(47)) * np.cos(lat) * np.cos(lon - np.radians(8)), -1, 1))
    assert (a["cell"] == int(np.argmin(d))).all()
