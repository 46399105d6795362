// ref_navigate -- runs the reference's Navigate action, unmodified and
// compiled from the reference tree, on a small scene, and prints what it
// builds and decides (TEST INFRASTRUCTURE ONLY: the tests pin the oracle's
// Navigate restatement against it; nothing in the product links or runs it).
//
// The scene: a non-periodic W x H lattice (as ref_harness builds it) with a
// Geography (altitude, ice), a Navigation (Navigation::setData / setBridges,
// QHG4/core/Navigation.cpp:24-60) and a tut_SexualPop population holding
// the listed agents (SPopulation::addAgent).  Navigate<tut_SexualAgent>
// (QHG4/actions/Navigate.cpp) gets its attributes, runs preLoop, and then
// execute for every agent in turn, the agent's life state (with
// LIFE_STATE_MOVING when listed as moving) set first.  A copy of the WELL512
// state taken before each call tells which draws the call took; the
// population's move list tells which moves it registered.  An optional
// geography event (new altitudes, then notify(EVENT_ID_GEO) and
// notify(EVENT_ID_FLUSH)) is followed by the tables, bridges and decisions
// again.
//
// usage: ref_navigate <scene.txt>
// scene: "grid W H", "alt" + W*H values, "ice" + W*H values,
//        "params decay dist0 prob0 bridge_prob", "ports K" then per port
//        "c n d1 dist1 ... dn distn", "bridges B" then B pairs, "seed S",
//        "agents M" then per agent "cell life moving", optionally
//        "geo" + W*H altitudes.
// output: "table c n+1 (target prob)*" per port table (probabilities as %a),
//         "bridge a b" per current bridge, "agent i cell ndraws draws*
//         nmoves dests*" per agent (draws as %a); then "geo" and the same
//         again when the scene has a geography event.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <omp.h>

// the population and the action in one translation unit, as the reference's
// population sources include their actions (both include clsutils.cpp,
// whose functions outside templates allow one copy per program)
#include "EventConsts.h"
#include "Geography.h"
#include "IDGen.h"
#include "MessLogger.h"
#include "Navigation.h"
#include "SCellGrid.h"
#include "WELL512.h"
#include "tut_SexualPop.cpp"
#include "Navigate.cpp"

namespace {

// Navigate with its protected state readable and its attributes settable
struct NavProbe : public Navigate<tut_SexualAgent> {
    NavProbe(SPopulation<tut_SexualAgent> *pop, SCellGrid *cg, WELL512 **w)
        : Navigate<tut_SexualAgent>(pop, cg, "", w) {}
    void set(double decay, double dist0, double prob0, double bridge) {
        this->m_dDecay = decay;
        this->m_dDist0 = dist0;
        this->m_dProb0 = prob0;
        this->m_dBridgeProb = bridge;
    }
    const jumpprobmap &jumps() const { return this->m_mJumpProbs; }
    const bridgelist &bridges() const { return this->m_vCurBridges; }
};

SCellGrid *lattice(int W, int H) {
    stringmap surf;
    surf[SURF_TYPE] = SURF_LATTICE;
    surf[SURF_LTC_LINKS] = "4";
    SCellGrid *cg = new SCellGrid(0, (uint)(W * H), surf);
    cg->m_aCells = new SCell[W * H];
    cg->m_iMaxNeighbors = 4;
    cg->m_iConnectivity = 4;
    for (int y = 0; y < H; y++) {
        for (int x = 0; x < W; x++) {
            const int i = y * W + x;
            SCell &c = cg->m_aCells[i];
            c.m_iGlobalID = i;
            c.m_iNumNeighbors = 0;
            for (int j = 0; j < MAX_NEIGH; j++) c.m_aNeighbors[j] = -1;
            cg->m_mIDIndexes[i] = i;
        }
    }
    return cg;
}

void print_state(NavProbe *nav) {
    for (const auto &e : nav->jumps()) {
        const int n = e.second.first + 1;
        std::printf("table %d %d", e.first, n);
        for (int j = 0; j < n; j++) std::printf(" %d %a", e.second.second[j].first, e.second.second[j].second);
        std::printf("\n");
    }
    for (const auto &b : nav->bridges()) std::printf("bridge %d %d\n", b.first, b.second);
}

void run_agents(NavProbe *nav, tut_SexualPop *pop, WELL512 *well, const std::vector<int> &cells,
                const std::vector<int> &life, const std::vector<int> &moving) {
    std::vector<int> *ml = pop->getMoveList()[0];
    for (size_t i = 0; i < cells.size(); i++) {
        tut_SexualAgent &a = pop->m_aAgents[(int)i];
        a.m_iCellIndex = cells[i];
        a.m_iLifeState = (uint)life[i] | (moving[i] ? LIFE_STATE_MOVING : 0u);
        uint32_t st[STATE_SIZE];
        std::memcpy(st, well->getState(), sizeof(st));
        WELL512 copy(st, well->getIndex());
        const unsigned long before = well->getCount();
        ml->clear();
        nav->execute((int)i, 0.0f);
        const unsigned long used = well->getCount() - before;
        std::printf("agent %zu %d %lu", i, cells[i], used);
        for (unsigned long k = 0; k < used; k++) std::printf(" %a", copy.wrandd());
        std::printf(" %zu", ml->size() / 3);
        for (size_t k = 0; k + 2 < ml->size(); k += 3) std::printf(" %d", (*ml)[k + 2]);
        std::printf("\n");
    }
    ml->clear();
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: %s <scene.txt>\n", argv[0]);
        return 2;
    }
    omp_set_num_threads(1);
    // the message log, as QHGMain creates it (Navigate logs a failing port)
    MessLogger::create(std::string(argv[1]) + ".log");
    std::ifstream in(argv[1]);
    int W = 0, H = 0;
    std::vector<double> alt, geo_alt;
    std::vector<int> ice, cells, life, moving;
    double decay = 0, dist0 = 0, prob0 = 0, bridge = 0;
    distancemap dests;
    bridgelist bridges;
    uint64_t seed = 1;
    for (std::string w; in >> w;) {
        if (w == "grid") {
            in >> W >> H;
        } else if (w == "alt" || w == "geo") {
            std::vector<double> &v = (w == "alt") ? alt : geo_alt;
            v.resize((size_t)(W * H));
            for (double &x : v) in >> x;
        } else if (w == "ice") {
            ice.resize((size_t)(W * H));
            for (int &x : ice) in >> x;
        } else if (w == "params") {
            in >> decay >> dist0 >> prob0 >> bridge;
        } else if (w == "ports") {
            int K = 0;
            in >> K;
            for (int k = 0; k < K; k++) {
                int c = 0, n = 0;
                in >> c >> n;
                distlist dl;
                for (int j = 0; j < n; j++) {
                    int d = 0;
                    double dist = 0;
                    in >> d >> dist;
                    dl[d] = dist;
                }
                dests[c] = dl;
            }
        } else if (w == "bridges") {
            int B = 0;
            in >> B;
            for (int b = 0; b < B; b++) {
                int x = 0, y = 0;
                in >> x >> y;
                bridges.push_back(bridgedef(x, y));
            }
        } else if (w == "seed") {
            in >> seed;
        } else if (w == "agents") {
            int M = 0;
            in >> M;
            cells.resize((size_t)M);
            life.resize((size_t)M);
            moving.resize((size_t)M);
            for (int i = 0; i < M; i++) in >> cells[(size_t)i] >> life[(size_t)i] >> moving[(size_t)i];
        }
    }
    if (W <= 0 || H <= 0 || alt.size() != (size_t)(W * H) || ice.size() != (size_t)(W * H)) {
        std::fprintf(stderr, "bad scene\n");
        return 2;
    }
    const int nc = W * H;
    SCellGrid *cg = lattice(W, H);
    Geography *geo = new Geography(cg, (uint)nc, 4, 6371.3);
    for (int c = 0; c < nc; c++) {
        geo->m_adAltitude[c] = alt[(size_t)c];
        geo->m_abIce[c] = ice[(size_t)c] != 0;
    }
    cg->setGeography(geo);
    Navigation *navi = new Navigation(cg);
    navi->setData(dests, 1.0);
    navi->setBridges(bridges);
    cg->setNavigation(navi);

    IDGen *idg[1] = {new IDGen(0, 0, 1)};
    uint32_t state[STATE_SIZE];
    for (unsigned k = 0; k < STATE_SIZE; k++) state[k] = (uint32_t)(seed * 2654435761u + 97u * k) | 1u;
    uint seeds[NUM_SEEDS] = {};
    tut_SexualPop *pop = new tut_SexualPop(cg, nullptr, 65536, idg, state, seeds);
    for (size_t i = 0; i < cells.size(); i++) {
        char line[128];
        std::snprintf(line, sizeof(line), "1 %zu 0.0 0 10.0 0.0", i + 1);
        if (pop->addAgent(cells[i], line, false) != 0) {
            std::fprintf(stderr, "addAgent failed\n");
            return 1;
        }
    }
    pop->updateTotal();
    uint32_t wst[STATE_SIZE];
    for (unsigned k = 0; k < STATE_SIZE; k++) wst[k] = (uint32_t)(seed * 40503u + 7919u * (k + 1)) | 1u;
    WELL512 *well = new WELL512(wst);
    WELL512 *wells[1] = {well};
    NavProbe *nav = new NavProbe(pop, cg, wells);
    nav->set(decay, dist0, prob0, bridge);
    if (nav->preLoop() != 0) {
        std::fprintf(stderr, "preLoop failed\n");
        return 1;
    }
    print_state(nav);
    run_agents(nav, pop, well, cells, life, moving);
    if (geo_alt.size() == (size_t)nc) {
        for (int c = 0; c < nc; c++) geo->m_adAltitude[c] = geo_alt[(size_t)c];
        nav->notify(nullptr, EVENT_ID_GEO, nullptr);
        nav->notify(nullptr, EVENT_ID_FLUSH, nullptr);
        std::printf("geo\n");
        print_state(nav);
        run_agents(nav, pop, well, cells, life, moving);
    }
    std::fflush(stdout);
    return 0;
}
