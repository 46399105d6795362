// plugin_host -- the reference's executable side around a population
// plugin, compiled from the reference tree (TEST INFRASTRUCTURE ONLY: it
// drives the MI355X PopBase plugin the way QHG4's application does, so the
// tests can compare what the plugin computes with the oracle).
//
// Every call on the population goes through the reference's own code:
//   DynPopFactory     scans <plugin dir> for *Wrapper.so, dlopens the one
//                     whose getInfo() names <class>, calls its createPop
//                     (QHG4/populations/DynPopFactory.cpp:80-190), then
//                     readSpeciesData with the species XML read by
//                     ParamProvider2 (QHG4/io/ParamProvider2.cpp)
//   addAgent          per agent data line, then updateTotal and
//                     updateNumAgentsPerCell, as SimParams::readAgentData
//                     does (QHG4/app/SimParams.cpp:2155-2245)
//   PopLooper         addPop (setPrioList, getPrios), preLoop, doStep
//                     (QHG4/core/PopLooper.cpp:96-202), postLoop
//   events            updateEvent then flushEvents on the population, with
//                     the Geography arrays rewritten in place first, as
//                     Simulator::runLoop and the geography reader do
//                     (QHG4/app/Simulator.cpp:335-374)
//   PopWriter         write / dump (QHG4/io/PopWriter.cpp:40-150)
//   PopReader         restore of a dump into a fresh plugin population
//                     (QHG4/io/PopReader.cpp:133-183)
// The grid is an SCellGrid with Geography and Vegetation objects of the
// reference, filled from raw arrays.
//
// usage: plugin_host <plugin dir> <class> <species.xml> <dir>
//   <dir>/grid.bin      int32 num_cells, int32 connectivity,
//                       int32 neighbors[num_cells*connectivity], uint8 counts[num_cells]
//   <dir>/<field>.f64   altitude, longitude, latitude, base_npp (raw doubles)
//   <dir>/agents.txt    "<cell>\t<agent data line>"
//   <dir>/plan.txt      "seed <n>", "steps <n>", "geo <step> <altitude file> <ice file>",
//                       "dump <step>" (dump there, restore into a new population
//                       through PopReader and go on with that one),
//                       "nav <file>" (the grid's Navigation, as NavGroupReader
//                       fills it: Navigation::setData / setBridges), "navev
//                       <step> <file>" (the Navigation replaced before that
//                       step, then updateEvent(EVENT_ID_NAV) and flushEvents)
//   navigation file     "ports K", K lines "cell n dest dist ...", "bridges B",
//                       B lines "cell cell"
// writes <dir>/out.qdf (PopWriter, after the last step), <dir>/snap.txt (the
// per-agent accessors and per-cell counts) and <dir>/log.txt.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "ArrayShare.h"
#include "DynPopFactory.h"
#include "EventConsts.h"
#include "Geography.h"
#include "IDGen.h"
#include "Navigation.h"
#include "ParamProvider2.h"
#include "PopBase.h"
#include "PopLooper.h"
#include "PopReader.h"
#include "PopWriter.h"
#include "SCellGrid.h"
#include "Vegetation.h"
#include "WELL512.h"

namespace {

template <typename T>
bool read_raw(const std::string &path, T *out, size_t n) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    const size_t got = std::fread(out, sizeof(T), n, f);
    std::fclose(f);
    return got == n;
}

int die(const char *what, const std::string &detail) {
    std::fprintf(stderr, "plugin_host: %s %s\n", what, detail.c_str());
    return 1;
}

// a navigation file into the grid's Navigation (setData, setBridges)
bool read_navigation(const std::string &path, Navigation *nav) {
    std::ifstream in(path);
    distancemap dests;
    bridgelist bridges;
    std::string w;
    int K = 0, B = 0;
    if (!(in >> w) || w != "ports" || !(in >> K)) return false;
    for (int k = 0; k < K; k++) {
        int c = 0, n = 0;
        if (!(in >> c >> n)) return false;
        distlist dl;
        for (int j = 0; j < n; j++) {
            int d = 0;
            double dist = 0;
            if (!(in >> d >> dist)) return false;
            dl[d] = dist;
        }
        dests[c] = dl;
    }
    if (!(in >> w) || w != "bridges" || !(in >> B)) return false;
    for (int b = 0; b < B; b++) {
        int x = 0, y = 0;
        if (!(in >> x >> y)) return false;
        bridges.push_back(bridgedef(x, y));
    }
    nav->setData(dests, 1.0);
    nav->setBridges(bridges);
    return true;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 5) return die("usage:", "plugin_host <plugin dir> <class> <species.xml> <dir>");
    const std::string sodir = argv[1], cls = argv[2], xml = argv[3], dir = argv[4];
    std::setvbuf(stdout, nullptr, _IOLBF, 0);

    // grid: cell i has global id i
    int32_t hdr[2];
    FILE *g = std::fopen((dir + "/grid.bin").c_str(), "rb");
    if (!g || std::fread(hdr, 4, 2, g) != 2) return die("cannot read", dir + "/grid.bin");
    const int nc = hdr[0], conn = hdr[1];
    if (conn > MAX_NEIGH) return die("too many neighbours in", dir + "/grid.bin");
    std::vector<int32_t> nbr((size_t)nc * conn);
    std::vector<uint8_t> nn((size_t)nc);
    if (std::fread(nbr.data(), 4, nbr.size(), g) != nbr.size() || std::fread(nn.data(), 1, nn.size(), g) != nn.size())
        return die("short grid file", dir);
    std::fclose(g);
    stringmap surf;
    surf[SURF_TYPE] = SURF_EQSAHEDRON;
    SCellGrid *cg = new SCellGrid(0, (uint)nc, surf);
    cg->m_aCells = new SCell[nc];
    cg->m_iConnectivity = conn;
    cg->m_iMaxNeighbors = conn;
    for (int c = 0; c < nc; c++) {
        SCell &cell = cg->m_aCells[c];
        cell.m_iGlobalID = c;
        cell.m_iNumNeighbors = nn[(size_t)c];
        for (int j = 0; j < MAX_NEIGH; j++) cell.m_aNeighbors[j] = (j < conn) ? nbr[(size_t)c * conn + j] : -1;
        cg->m_mIDIndexes[c] = c;
    }
    Geography *geo = new Geography(cg, (uint)nc, (uint)conn, 6371.3);
    Vegetation *veg = new Vegetation(cg, (uint)nc, 0);
    if (!read_raw(dir + "/altitude.f64", geo->m_adAltitude, (size_t)nc) ||
        !read_raw(dir + "/longitude.f64", geo->m_adLongitude, (size_t)nc) ||
        !read_raw(dir + "/latitude.f64", geo->m_adLatitude, (size_t)nc) ||
        !read_raw(dir + "/base_npp.f64", veg->m_adBaseANPP, (size_t)nc))
        return die("cannot read the environment in", dir);
    cg->setGeography(geo);
    cg->setVegetation(veg);

    // plan
    uint64_t seed = 0;
    int steps = 0, geo_step = -1, dump_step = -1, nav_step = -1;
    std::string geo_alt, geo_ice, nav_file, nav_file2;
    std::ifstream pl(dir + "/plan.txt");
    for (std::string w; pl >> w;) {
        if (w == "seed") pl >> seed;
        else if (w == "steps") pl >> steps;
        else if (w == "geo") pl >> geo_step >> geo_alt >> geo_ice;
        else if (w == "dump") pl >> dump_step;
        else if (w == "nav") pl >> nav_file;
        else if (w == "navev") pl >> nav_step >> nav_file2;
    }
    if (!nav_file.empty()) {
        Navigation *navi = new Navigation(cg);
        if (!read_navigation(dir + "/" + nav_file, navi)) return die("cannot read", nav_file);
        cg->setNavigation(navi);
    }

    // the application's WELL seed state (the plugin keys its draws on its
    // first two words) and id generators
    uint32_t state[STATE_SIZE] = {};
    state[0] = (uint32_t)seed;
    state[1] = (uint32_t)(seed >> 32);
    IDGen *idg[1] = {new IDGen(0, 0, 1)};

    FILE *log = std::fopen((dir + "/log.txt").c_str(), "w");
    if (!log) return die("cannot write", dir + "/log.txt");
    std::setvbuf(log, nullptr, _IOLBF, 0);

    PopLooper *looper = new PopLooper();
    stringvec sodirs;
    sodirs.push_back(sodir);
    DynPopFactory *fac = DynPopFactory::createInstance(sodirs, cg, looper, 65536, idg, state, nullptr);
    if (!fac) return die("no plugins in", sodir);
    ParamProvider2 *pp = ParamProvider2::createInstance(xml);
    if (!pp) return die("cannot read", xml);
    if (pp->selectClass(cls) != 0) return die("no class in the XML:", cls);
    PopBase *pop = fac->readPopulation(pp);
    if (!pop) return die("readPopulation failed for", cls);
    const std::string species = pop->getSpeciesName();
    std::fprintf(log, "created %s species %s\n", pop->getClassName().c_str(), species.c_str());

    std::ifstream ag(dir + "/agents.txt");
    size_t added = 0;
    for (std::string line; std::getline(ag, line);) {
        const size_t tab = line.find('\t');
        if (tab == std::string::npos) continue;
        std::vector<char> data(line.begin() + (long)tab + 1, line.end());
        data.push_back('\0');
        if (pop->addAgent(std::atoi(line.substr(0, tab).c_str()), data.data(), false) != 0)
            return die("addAgent failed on", line);
        added++;
    }
    pop->updateTotal();
    pop->updateNumAgentsPerCell();
    if (looper->addPop(pop) != 0) return die("addPop failed", "");
    if (looper->preLoop() != 0) return die("preLoop failed", "");
    std::fprintf(log, "added %zu agents, %lu effective\n", added, pop->getNumAgentsEffective());

    for (int s = 0; s < steps; s++) {
        const float t = (float)s;
        if (looper->doStep(t) != 0) return die("doStep failed at step", std::to_string(s));
        std::fprintf(log, "step %d alive %lu\n", s, pop->getNumAgentsEffective());
        if (s + 1 == geo_step) {
            std::vector<uint8_t> ice((size_t)nc);
            if (!read_raw(dir + "/" + geo_alt, geo->m_adAltitude, (size_t)nc) ||
                !read_raw(dir + "/" + geo_ice, ice.data(), (size_t)nc))
                return die("cannot read the event's arrays in", dir);
            for (int c = 0; c < nc; c++) geo->m_abIce[c] = ice[(size_t)c] != 0;
            char ev[] = "geo";
            pop->updateEvent(EVENT_ID_GEO, ev, (float)(s + 1));
            pop->flushEvents((float)(s + 1));
            std::fprintf(log, "geo event before step %d alive %lu\n", s + 1, pop->getNumAgentsEffective());
        }
        if (s + 1 == nav_step) {
            if (!cg->m_pNavigation || !read_navigation(dir + "/" + nav_file2, cg->m_pNavigation))
                return die("cannot read", nav_file2);
            char ev[] = "nav";
            pop->updateEvent(EVENT_ID_NAV, ev, (float)(s + 1));
            pop->flushEvents((float)(s + 1));
            std::fprintf(log, "nav event before step %d alive %lu\n", s + 1, pop->getNumAgentsEffective());
        }
        if (s + 1 == dump_step) {
            const std::string f = dir + "/dump.qdf";
            std::vector<PopBase *> v(1, pop);
            PopWriter *pw = new PopWriter(v);
            const int rc = pw->write(f, s + 1, 0.0f, "plugin_host dump", species, PW_ALL, 0);
            delete pw;
            if (rc != 0) return die("dump failed", f);
            PopBase *pop2 = fac->createPopulationByName(cls);
            if (!pop2) return die("createPopulationByName failed for", cls);
            PopReader *pr = PopReader::create(f);
            if (!pr) return die("cannot open", f);
            const int rr = pr->read(pop2, species, nc, true);
            delete pr;
            if (rr != 0) return die("restore failed", std::to_string(rr));
            looper->postLoop();
            delete looper;   // deletes the populations it holds (PopLooper.cpp:25-34)
            pop = pop2;
            looper = new PopLooper();
            if (looper->addPop(pop) != 0 || looper->preLoop() != 0) return die("restored population", "not started");
            std::fprintf(log, "dumped and restored after step %d alive %lu\n", s + 1, pop->getNumAgentsEffective());
        }
    }
    if (looper->postLoop() != 0) return die("postLoop failed", "");

    // the per-agent accessors and per-cell counts
    FILE *sn = std::fopen((dir + "/snap.txt").c_str(), "w");
    if (!sn) return die("cannot write", dir + "/snap.txt");
    pop->updateNumAgentsPerCell();
    const int first = pop->getFirstAgentIndex(), last = pop->getLastAgentIndex();
    std::fprintf(sn, "agents %d %d %lu %lu\n", first, last, pop->getNumAgentsEffective(), (ulong)pop->getMaxLoadedID());
    for (int i = first; i >= 0 && i <= last; i++)
        std::fprintf(sn, "%ld %d %ld %u %u %.9g\n", (long)pop->getAgentID(i), pop->getAgentCellIndex(i),
                     (long)pop->getAgentCellID(i), pop->getAgentLifeState(i), (unsigned)pop->getAgentGender(i),
                     (double)pop->getAgentBirthTime(i));
    for (int c = 0; c < nc; c++) std::fprintf(sn, "cell %d %lu\n", c, pop->getNumAgents(c));
    std::fclose(sn);

    std::vector<PopBase *> v(1, pop);
    PopWriter *pw = new PopWriter(v);
    const int rc = pw->write(dir + "/out.qdf", steps, 0.0f, "plugin_host", species, PW_ALL);
    delete pw;
    if (rc != 0) return die("write failed", dir + "/out.qdf");
    std::fprintf(log, "done agents %lu\n", pop->getNumAgentsEffective());
    std::fclose(log);
    delete looper;   // and with it the population, before its library is closed
    delete fac;
    return 0;
}
