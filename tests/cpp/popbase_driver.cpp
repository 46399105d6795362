// A minimal host driver over qhg::Population (include/qhg_population.hpp):
// it configures a species from module attributes, adds agents from agent
// data lines, and steps in PopLooper::doStep order
// (QHG4/core/PopLooper.cpp:166-202: initializeStep, doActions per
// priority, finalizeStep), with Simulator::runLoop's events between steps
// (QHG4/app/Simulator.cpp:335-374: updateEvent, then flushEvents).
//
// usage: popbase_driver <dir>
//   <dir>/grid.bin      int32 num_cells, int32 connectivity,
//                       int32 neighbors[num_cells*connectivity], uint8 counts[num_cells]
//   <dir>/<field>.f64   environment arrays (altitude, longitude, latitude,
//                       base_npp), raw doubles, each optional
//   <dir>/species.txt   tab-separated lines: class, species, seed, module,
//                       param <key> <value> (to the last module), prio <name> <value>
//   <dir>/agents.txt    "<cell>\t<agent data line>"
//   <dir>/plan.txt      "steps <n>" and "geo <step> <altitude file> <ice file>"
// writes <dir>/out.bin (the live agents, id order) and <dir>/log.txt
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "qhg_population.hpp"

namespace {

template <typename T>
bool read_raw(const std::string &path, std::vector<T> &out, size_t n) {
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f) return false;
    out.resize(n);
    const size_t got = std::fread(out.data(), sizeof(T), n, f);
    std::fclose(f);
    return got == n;
}

std::vector<std::string> split_tabs(const std::string &line) {
    std::vector<std::string> v;
    std::stringstream ss(line);
    std::string item;
    while (std::getline(ss, item, '\t')) v.push_back(item);
    return v;
}

int die(const char *what, const std::string &detail) {
    std::fprintf(stderr, "popbase_driver: %s %s\n", what, detail.c_str());
    return 1;
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) return die("usage:", "popbase_driver <dir>");
    const std::string dir = argv[1];
    // grid
    int32_t hdr[2];
    FILE *g = std::fopen((dir + "/grid.bin").c_str(), "rb");
    if (!g || std::fread(hdr, 4, 2, g) != 2) return die("cannot read", dir + "/grid.bin");
    const int32_t nc = hdr[0], conn = hdr[1];
    std::vector<int32_t> nbr((size_t)nc * conn);
    std::vector<uint8_t> nn((size_t)nc);
    if (std::fread(nbr.data(), 4, nbr.size(), g) != nbr.size() || std::fread(nn.data(), 1, nn.size(), g) != nn.size())
        return die("short grid file", dir);
    std::fclose(g);
    qhg_grid_desc grid{nc, conn, nbr.data(), nn.data(), nullptr};
    // environment
    std::vector<double> alt, lon, lat, npp;
    std::vector<uint8_t> ice;
    qhg_env_desc env{};
    if (read_raw(dir + "/altitude.f64", alt, (size_t)nc)) env.altitude = alt.data();
    if (read_raw(dir + "/longitude.f64", lon, (size_t)nc)) env.longitude = lon.data();
    if (read_raw(dir + "/latitude.f64", lat, (size_t)nc)) env.latitude = lat.data();
    if (read_raw(dir + "/base_npp.f64", npp, (size_t)nc)) env.base_npp = npp.data();
    // species
    std::string cls, species;
    uint64_t seed = 0;
    qhg::ModuleList mods;
    qhg::AttrMap prios;
    std::ifstream sp(dir + "/species.txt");
    for (std::string line; std::getline(sp, line);) {
        const auto f = split_tabs(line);
        if (f.size() >= 2 && f[0] == "class") cls = f[1];
        else if (f.size() >= 2 && f[0] == "species") species = f[1];
        else if (f.size() >= 2 && f[0] == "seed") seed = std::strtoull(f[1].c_str(), nullptr, 10);
        else if (f.size() >= 2 && f[0] == "module") mods.push_back({f[1], {}});
        else if (f.size() >= 3 && f[0] == "param" && !mods.empty()) mods.back().second[f[1]] = f[2];
        else if (f.size() >= 3 && f[0] == "prio") prios[f[1]] = f[2];
    }
    qhg::Population pop(grid, env, cls, species);
    if (pop.readSpeciesData(mods, prios, seed)) return die("readSpeciesData:", pop.lastError());
    if (pop.setPrioList()) return die("setPrioList:", pop.lastError());
    std::set<unsigned> levels;
    pop.getPrios(levels);
    // agents
    std::ifstream ag(dir + "/agents.txt");
    for (std::string line; std::getline(ag, line);) {
        const auto f = split_tabs(line);
        if (f.size() != 2) continue;
        if (pop.addAgent(std::atoi(f[0].c_str()), f[1].c_str())) return die("addAgent:", pop.lastError());
    }
    // plan
    int steps = 0, geo_step = -1;
    std::string geo_alt, geo_ice;
    std::ifstream pl(dir + "/plan.txt");
    for (std::string w; pl >> w;) {
        if (w == "steps") pl >> steps;
        else if (w == "geo") pl >> geo_step >> geo_alt >> geo_ice;
    }
    FILE *log = std::fopen((dir + "/log.txt").c_str(), "w");
    if (!log) return die("cannot write", dir + "/log.txt");
    if (pop.preLoop()) return die("preLoop:", pop.lastError());
    std::fprintf(log, "prios %zu agents %lu\n", levels.size(), pop.getNumAgentsEffective());
    for (int s = 0; s < steps; s++) {
        const float t = (float)s;
        if (pop.initializeStep(t)) return die("initializeStep:", pop.lastError());
        for (unsigned prio : levels)
            if (pop.doActions(prio, t)) return die("doActions:", pop.lastError());
        if (pop.finalizeStep()) return die("finalizeStep:", pop.lastError());
        std::fprintf(log, "step %d alive %lu\n", s, pop.getNumAgentsEffective());
        if (s + 1 == geo_step) {
            if (!read_raw(dir + "/" + geo_alt, alt, (size_t)nc) || !read_raw(dir + "/" + geo_ice, ice, (size_t)nc))
                return die("cannot read the event's arrays in", dir);
            env.altitude = alt.data();
            env.ice = ice.data();
            pop.setEnvironment(env);
            pop.updateEvent(qhg::kEventIdGeo, (float)(s + 1));
            if (pop.flushEvents((float)(s + 1))) return die("flushEvents:", pop.lastError());
            std::fprintf(log, "geo event before step %d alive %lu\n", s + 1, pop.getNumAgentsEffective());
        }
    }
    if (pop.postLoop()) return die("postLoop:", pop.lastError());
    qhg::Agents a;
    if (pop.download(a)) return die("download:", pop.lastError());
    FILE *o = std::fopen((dir + "/out.bin").c_str(), "wb");
    if (!o) return die("cannot write", dir + "/out.bin");
    const int64_t n = (int64_t)a.size();
    std::fwrite(&n, 8, 1, o);
    std::fwrite(a.id.data(), 8, a.size(), o);
    std::fwrite(a.cell.data(), 4, a.size(), o);
    std::fwrite(a.life.data(), 1, a.size(), o);
    std::fwrite(a.gender.data(), 1, a.size(), o);
    std::fwrite(a.birth_time.data(), 8, a.size(), o);
    std::fwrite(a.age.data(), 8, a.size(), o);
    std::fwrite(a.last_birth.data(), 8, a.size(), o);
    std::fwrite(a.mate_id.data(), 8, a.size(), o);
    std::fwrite(a.num_babies.data(), 4, a.size(), o);
    std::fclose(o);
    std::fprintf(log, "done agents %lld max_id %ld\n", (long long)n, pop.getMaxLoadedID());
    std::fclose(log);
    return 0;
}
