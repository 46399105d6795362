// Host-only checks of include/qhg_population.hpp (no GPU calls): agent data
// lines, the parameter block from module attributes, priorities, and the
// errors of a population that has no device yet.  Prints "ok" and returns 0.
#include <cmath>
#include <cstdio>
#include <cstring>

#include "qhg_population.hpp"

#define CHECK(c)                                                                  \
    do {                                                                          \
        if (!(c)) {                                                               \
            std::fprintf(stderr, "%s:%d: check failed: %s\n", __FILE__, __LINE__, #c); \
            return 1;                                                             \
        }                                                                         \
    } while (0)

int main() {
    // agent data lines: LifeState AgentID BirthTime Gender [Age [LastBirth]]
    double v[6];
    int k = 0;
    CHECK(qhg::parse_agent_data("1 17 -3.5 0 3.5 -1", v, &k) == 0 && k == 6);
    CHECK(v[0] == 1 && v[1] == 17 && v[2] == -3.5 && v[3] == 0 && v[4] == 3.5 && v[5] == -1);
    CHECK(qhg::parse_agent_data("5;42;0.25;1", v, &k) == 0 && k == 4 && v[1] == 42 && v[3] == 1);
    CHECK(qhg::parse_agent_data("1,2:3 4", v, &k) == 0 && k == 4);
    CHECK(qhg::parse_agent_data("1 2 3", v, &k) != 0);          // too few fields
    CHECK(qhg::parse_agent_data("1 two 3 4", v, &k) != 0);      // not a number
    CHECK(qhg::parse_agent_data("", v, &k) != 0);

    // the tutorial's tut_SexualPop modules (QHG4/tutorial_data/xmldat/tut_Sexual.xml)
    qhg::ModuleList mods = {
        {"GetOld", {}},
        {"ATanDeath", {{"ATanDeath_max_age", "70"}, {"ATanDeath_range", "15"}, {"ATanDeath_slope", "0.4"}}},
        {"RandomMove", {{"RandomMove_prob", "0.07"}}},
        {"Fertility", {{"Fertility_min_age", "15"}, {"Fertility_max_age", "50"}, {"Fertility_interbirth", "2"}}},
        {"Verhulst", {{"Verhulst_b0", "0.8"}, {"Verhulst_d0", "0.001"}, {"Verhulst_theta", "0.1"}, {"Verhulst_K", "10"}}},
        {"RandomPair", {}},
    };
    qhg::AttrMap prios = {{"GetOld", "1"}, {"ATanDeath", "2"}, {"RandomMove", "3"},
                          {"Fertility", "4"}, {"RandomPair", "5"}, {"Verhulst", "6"}};
    qhg_params p;
    CHECK(qhg::params_from_modules("tut_SexualPop", mods, prios, 7, &p) == 0);
    CHECK(p.num_actions == 6 && p.seed == 7);
    CHECK(p.verh_b0 == 0.8 && p.rmove_prob == 0.07 && p.fert_interbirth == 2 && p.atan_slope == 0.4);
    double val = 0;
    CHECK(qhg_params_get_attribute(&p, "Verhulst_K", &val) == 0 && val == 10);
    // a malformed priority and a missing attribute are refused
    qhg::AttrMap bad = prios;
    bad["GetOld"] = "first";
    CHECK(qhg::params_from_modules("tut_SexualPop", mods, bad, 7, &p) != 0);
    qhg::ModuleList short_mods = mods;
    short_mods[4].second.erase("Verhulst_K");
    CHECK(qhg::params_from_modules("tut_SexualPop", short_mods, prios, 7, &p) != 0);
    CHECK(std::strstr(qhg_gpu_last_error(), "Verhulst_K") != nullptr);

    // a population before readSpeciesData: every call reports the missing device
    int32_t nbr[4] = {-1, -1, -1, -1};
    uint8_t nn[1] = {0};
    qhg_grid_desc grid{1, 4, nbr, nn, nullptr};
    qhg_env_desc env{};
    qhg::Population pop(grid, env, "tut_SexualPop", "sapiens");
    CHECK(pop.getClassName() == "tut_SexualPop" && pop.getSpeciesName() == "sapiens");
    CHECK(pop.preLoop() == QHG_ERR_STATE);
    CHECK(pop.lastError().find("readSpeciesData") != std::string::npos);
    CHECK(pop.initializeStep(0.f) == QHG_ERR_STATE);
    CHECK(pop.getNumAgentsEffective() == 0);
    CHECK(pop.getMaxLoadedID() == -1);
    CHECK(pop.updateEvent(qhg::kEventIdGeo, 1.f) == 0);
    CHECK(pop.flushEvents(1.f) == 0);   // nothing to apply without a device
    std::set<unsigned> levels;
    CHECK(pop.getPrios(levels) == 0 && levels.empty());
    std::puts("ok");
    return 0;
}
