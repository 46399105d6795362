// Host-only exercise of the oracle (oracle/qhg_oracle.c), the species
// configuration (qhg4_amd/csrc/config.cpp) and the QDF I/O
// (qhg4_amd/csrc/qdf_io.cpp), built with -fsanitize=address,undefined by
// tests/cpp/build_driver.py (build_sanitize) and run by
// tests/test_sanitize.py.  Every step of every action set runs with several
// OpenMP threads; the configuration and QDF paths run their error cases too.
// Any sanitizer report aborts the program (the build sets
// -fno-sanitize-recover=all).
//
// The library entry points that need a GPU (population upload/download and
// clock) are defined below to refuse with an error: this build has no
// device, and nothing here calls them.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "qhg_gpu.h"
#include "qhg_qdf.h"
extern "C" {
#include "qhg_oracle.h"
}

// -- the host pieces of libqhg4_amd that config.cpp and qdf_io.cpp use ------
namespace {
thread_local std::string g_err;
}
int qhg_internal_fail(int code, const std::string &msg) {
    g_err = msg;
    return code;
}
extern "C" {
const char *qhg_gpu_last_error(void) { return g_err.c_str(); }
int qhg_gpu_genome_words(int32_t genome_size) { return (genome_size + 31) / 32; }
int qhg_gpu_download_agents(qhg_pop *, qhg_agents *) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
int qhg_gpu_upload_agents(qhg_pop *, const qhg_agents *) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
int qhg_gpu_get_clock(qhg_pop *, int64_t *, double *, int64_t *) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
int qhg_gpu_set_clock(qhg_pop *, int64_t, double, int64_t) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
int qhg_gpu_get_params(qhg_pop *, qhg_params *) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
int qhg_gpu_num_alive(qhg_pop *, int64_t *) { return qhg_internal_fail(QHG_ERR_STATE, "no device"); }
}

namespace {

int failures = 0;
#define CHECK(cond)                                                                  \
    do {                                                                             \
        if (!(cond)) {                                                               \
            std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #cond);  \
            failures++;                                                              \
        }                                                                            \
    } while (0)

uint64_t splitmix(uint64_t &s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

struct World {
    int W, H, nc;
    std::vector<int32_t> nbr;
    std::vector<uint8_t> nn, ice, coastal;
    std::vector<double> alt, lon, lat, water, npp;
    qhg_grid_desc grid{};
    qhg_env_desc env{};
};

// a non-periodic lattice in GridFactory's slot order (+x, +y, -x, -y; -1 at
// the border) with random environment arrays
World make_world(int W, int H, uint64_t seed) {
    World w;
    w.W = W;
    w.H = H;
    w.nc = W * H;
    w.nbr.assign((size_t)w.nc * 4, -1);
    w.nn.assign((size_t)w.nc, 0);
    for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++) {
            const int c = y * W + x;
            int32_t *r = &w.nbr[(size_t)c * 4];
            r[0] = (x + 1 < W) ? c + 1 : -1;
            r[1] = (y + 1 < H) ? c + W : -1;
            r[2] = (x > 0) ? c - 1 : -1;
            r[3] = (y > 0) ? c - W : -1;
            for (int k = 0; k < 4; k++) w.nn[(size_t)c] += (r[k] >= 0) ? 1 : 0;
        }
    uint64_t s = seed;
    w.alt.resize((size_t)w.nc);
    w.lon.resize((size_t)w.nc);
    w.lat.resize((size_t)w.nc);
    w.water.resize((size_t)w.nc);
    w.npp.resize((size_t)w.nc);
    w.ice.resize((size_t)w.nc);
    w.coastal.resize((size_t)w.nc);
    for (int c = 0; c < w.nc; c++) {
        const double u = (double)(splitmix(s) >> 11) / 9007199254740992.0;
        w.alt[(size_t)c] = -500.0 + 3000.0 * u;
        w.lon[(size_t)c] = 100.0 + 0.1 * (c % W);
        w.lat[(size_t)c] = -20.0 + 0.1 * (c / W);
        w.water[(size_t)c] = (u > 0.9) ? 1.0 : 0.0;
        w.npp[(size_t)c] = 1.4 * u;
        w.ice[(size_t)c] = (u < 0.02) ? 1 : 0;
        w.coastal[(size_t)c] = (u > 0.8) ? 1 : 0;
    }
    w.grid = qhg_grid_desc{w.nc, 4, w.nbr.data(), w.nn.data(), nullptr};
    w.env = qhg_env_desc{w.alt.data(), w.lon.data(), w.lat.data(), w.water.data(), w.ice.data(),
                         w.coastal.data(), w.npp.data(), nullptr, nullptr};
    return w;
}

struct Pop {
    std::vector<int64_t> id, mate;
    std::vector<int32_t> cell, nbab;
    std::vector<uint8_t> life, gender;
    std::vector<double> birth, age, lastb;
    std::vector<uint64_t> genome;
    qhg_agents view(int gw) {
        return qhg_agents{(int64_t)id.size(), id.data(), cell.data(), life.data(), gender.data(), birth.data(),
                          age.data(), lastb.data(), mate.data(), nbab.data(), gw ? genome.data() : nullptr};
    }
};

Pop make_agents(int n, int nc, int gw, uint64_t seed) {
    Pop p;
    uint64_t s = seed;
    std::vector<uint64_t> base((size_t)gw);
    for (auto &b : base) b = splitmix(s);
    for (int i = 0; i < n; i++) {
        p.id.push_back(i + 1);
        p.cell.push_back((int32_t)(splitmix(s) % (uint64_t)nc));
        p.life.push_back(QHG_LIFE_ALIVE);
        p.gender.push_back((uint8_t)(splitmix(s) & 1));
        const double a = (double)(splitmix(s) % 40);
        p.birth.push_back(-a);
        p.age.push_back(a);
        p.lastb.push_back(-1.0);
        p.mate.push_back(-3);
        p.nbab.push_back(0);
        for (int k = 0; k < gw; k++) {
            uint64_t v = base[(size_t)k];
            if ((splitmix(s) & 7) == 0) v ^= 1ull << (splitmix(s) % 64);
            p.genome.push_back(v);
        }
    }
    return p;
}

void add_action(qhg_params &P, int kind, int prio, const char *name) {
    const int i = P.num_actions++;
    P.action_kind[i] = kind;
    P.action_prio[i] = prio;
    std::snprintf(P.action_name[i], QHG_NAME_LEN, "%s", name);
}

qhg_params base_params(uint64_t seed) {
    qhg_params P;
    std::memset(&P, 0, sizeof P);
    P.seed = seed;
    P.fert_min_age = 15;
    P.fert_max_age = 45;
    P.fert_interbirth = 3;
    P.verh_b0 = 0.8;
    P.verh_d0 = 0.001;
    P.verh_theta = 0.1;
    P.verh_K = 30;
    P.rmove_prob = 0.1;
    P.wmove_prob = 0.2;
    P.npp_min = 0;
    P.npp_max = 1.5;
    P.npp_k_min = 0;
    P.npp_k_max = 60;
    P.npp_efficiency = 1.0;
    P.npp_water_factor = 0.5;
    P.npp_coastal_factor = 0.7;
    P.npp_coastal_min_lat = -30;
    P.npp_coastal_max_lat = 30;
    P.oad_max_age = 60;
    P.oad_uncertainty = 0.1;
    P.atan_max_age = 60;
    P.atan_range = 10;
    P.atan_slope = 0.2;
    return P;
}

// one model on the oracle: steps with 1 and 4 threads must agree, an
// environment event and a parameter change in between
void run_model(const char *what, const World &w, qhg_params P, int n, int gw, int steps,
               const qhg_nav_desc *nav = nullptr) {
    Pop a = make_agents(n, w.nc, gw, P.seed);
    std::vector<int64_t> alive;
    for (int th : {1, 4}) {
        orc_set_threads(th);
        orc *o = nullptr;
        int rc = orc_create(&w.grid, &w.env, &P, &o);
        CHECK(rc == 0 && o);
        if (rc || !o) return;
        qhg_agents v = a.view(gw);
        CHECK(orc_set_agents(o, &v) == 0);
        if (nav) CHECK(orc_set_navigation(o, nav) == 0);
        CHECK(orc_pre_loop(o) == 0);
        orc_run(o, steps / 2);
        std::vector<double> alt2(w.alt);
        for (int c = 0; c < w.nc; c += 7) alt2[(size_t)c] = -10.0;
        qhg_env_desc e2 = w.env;
        e2.altitude = alt2.data();
        orc_env_event(o, &e2, 1);
        qhg_params P2 = P;
        P2.verh_b0 *= 0.9;
        CHECK(orc_set_params(o, &P2) == 0);
        if (nav) {
            // a navigation event: the same sea-ways without the last port
            qhg_nav_desc n2 = *nav;
            n2.num_ports = nav->num_ports > 0 ? nav->num_ports - 1 : 0;
            CHECK(orc_set_navigation(o, &n2) == 0);
        }
        orc_run(o, steps - steps / 2);
        qhg_stats st{};
        orc_stats(o, &st);
        std::vector<uint32_t> cc((size_t)w.nc);
        orc_cell_counts(o, cc.data());
        int64_t sum = 0;
        for (uint32_t x : cc) sum += x;
        CHECK(sum == orc_num_alive(o));
        std::vector<double> caps((size_t)w.nc), wts((size_t)w.nc * 5);
        orc_capacities(o, caps.data());
        orc_env_weights(o, wts.data());
        const int64_t m = orc_num_alive(o);
        Pop out;
        out.id.resize((size_t)m);
        out.cell.resize((size_t)m);
        out.life.resize((size_t)m);
        out.gender.resize((size_t)m);
        out.birth.resize((size_t)m);
        out.age.resize((size_t)m);
        out.lastb.resize((size_t)m);
        out.mate.resize((size_t)m);
        out.nbab.resize((size_t)m);
        out.genome.resize((size_t)(m * gw));
        qhg_agents ov = out.view(gw);
        CHECK(orc_get_agents(o, &ov) == 0);
        alive.push_back(m);
        orc_destroy(o);
    }
    CHECK(alive.size() == 2 && alive[0] == alive[1]);
    std::printf("model %-28s alive %lld\n", what, alive.empty() ? -1LL : (long long)alive[0]);
}

void oracle_models() {
    const World w = make_world(24, 20, 7);
    {
        qhg_params P = base_params(11);
        add_action(P, QHG_ACT_GETOLD, 1, "GetOld");
        add_action(P, QHG_ACT_FERTILITY, 2, "Fertility");
        add_action(P, QHG_ACT_RANDOMPAIR, 3, "RandomPair");
        add_action(P, QHG_ACT_VERHULST, 4, "Verhulst");
        add_action(P, QHG_ACT_RANDOMMOVE, 5, "RandomMove");
        add_action(P, QHG_ACT_ATANDEATH, 6, "ATanDeath");
        run_model("tut_Sexual + ATanDeath", w, P, 3000, 0, 30);
    }
    for (int kind : {0, QHG_ACT_ASSORTPAIR, QHG_ACT_ASSORTPROT}) {
        qhg_params P = base_params(12 + kind);
        add_action(P, QHG_ACT_NPPCAP, 1, "NPPCapacity");
        add_action(P, QHG_ACT_GETOLD, 2, "GetOld");
        add_action(P, QHG_ACT_OLDAGEDEATH, 3, "OldAgeDeath");
        add_action(P, QHG_ACT_WEIGHTEDMOVE, 4, "WeightedMove");
        add_action(P, QHG_ACT_MULTIEVAL, 5, "MultiEvaluator[Alt+NPP]");
        add_action(P, QHG_ACT_FERTILITY, 6, "Fertility");
        add_action(P, kind ? kind : QHG_ACT_RANDOMPAIR, 7, "Pair");
        add_action(P, QHG_ACT_VERHULSTVARK, 8, "VerhulstVarK");
        add_action(P, QHG_ACT_GENETICS, 9, "Genetics");
        P.me_mode = QHG_ME_MUL_SIMPLE;
        P.me_num = 2;
        P.me_input[0] = QHG_EV_ALTITUDE;
        P.me_input[1] = QHG_EV_CAPACITY;
        P.me_weight[0] = P.me_weight[1] = 1.0;
        P.me_pl_n[0] = 3;
        const double xs[3] = {-500, 1000, 3000}, vs[3] = {1, 0.5, 0};
        for (int k = 0; k < 3; k++) {
            P.me_pl_x[0][k] = xs[k];
            P.me_pl_v[0][k] = vs[k];
        }
        P.gen_genome_size = 100;
        P.gen_num_crossover = 2;
        P.gen_mutation_rate = 0.01;
        P.assort_cutoff = (kind == QHG_ACT_ASSORTPROT) ? 1.0 : 0.9;
        P.assort_permute = 1;
        const char *names[] = {"OoA-style + genetics", "... + AssortativePairHamming", "... + AssortativePair"};
        run_model(names[kind == 0 ? 0 : (kind == QHG_ACT_ASSORTPAIR ? 1 : 2)], w, P, 3000, 2 * 4, 24);
        if (kind != 0) continue;
        // ... + Navigate: ports with 0-6 destinations (one whose
        // probabilities add up to more than 1), bridges (one from a cell to
        // itself, some across water), a navigation event midway
        add_action(P, QHG_ACT_NAVIGATE, 5, "Navigate");
        P.nav_decay = -0.001;
        P.nav_dist0 = 150;
        P.nav_prob0 = 0.2;
        P.nav_bridge_prob = 0.4;
        uint64_t s = 99;
        std::vector<int32_t> ports, off(1, 0), cells, bridges;
        std::vector<double> dists;
        for (int c = 3; c < w.nc; c += 11) {
            ports.push_back(c);
            const int k = (c == 47) ? 9 : (int)(splitmix(s) % 7);
            for (int j = 0; j < k; j++) {
                cells.push_back((c + 5 + 13 * j) % w.nc);
                dists.push_back((c == 47) ? 1.0 : 10.0 + (double)(splitmix(s) % 500));
            }
            off.push_back((int32_t)cells.size());
        }
        for (int b = 0; b < 12; b++) {
            bridges.push_back((int32_t)(splitmix(s) % (uint64_t)w.nc));
            bridges.push_back(b == 0 ? bridges.back() : (int32_t)(splitmix(s) % (uint64_t)w.nc));
        }
        const qhg_nav_desc nav{(int32_t)ports.size(), ports.data(), off.data(), cells.data(), dists.data(), 1.0,
                               12, bridges.data()};
        run_model("... + Navigate", w, P, 3000, 2 * 4, 24, &nav);
    }
}

void oracle_units() {
    // every MD5 padding case and the proteome reader at genome sizes that cut
    // strands at every word and codon offset
    std::vector<uint8_t> msg(300);
    uint64_t s = 5;
    for (auto &b : msg) b = (uint8_t)splitmix(s);
    uint8_t out[16];
    for (int n = 0; n <= 300; n++) orc_md5(msg.data(), n, out);
    for (int G : {1, 3, 4, 5, 31, 32, 33, 64, 65, 97, 256, 1100}) {
        const int gw = 2 * ((G + 31) / 32);
        std::vector<uint64_t> g((size_t)gw);
        for (auto &x : g) x = splitmix(s);
        const int mx = 2 * ((G / 3 + 1) / 2 + 1);
        std::vector<uint8_t> dg((size_t)mx * 16);
        const int n = orc_proteome(g.data(), gw, G, dg.data(), mx);
        CHECK(n >= 0 && n <= mx);
        CHECK(orc_prot_matches(dg.data(), n < mx ? n : mx, dg.data(), n < mx ? n : mx) >= (n < mx ? n : mx));
    }
    for (int b = 0; b < 200; b++) {
        std::vector<float> R(12 * 9);
        for (auto &r : R) r = (float)(splitmix(s) % 5) / 4.0f;
        std::vector<uint32_t> u(400);
        for (auto &x : u) x = (uint32_t)splitmix(s);
        std::vector<int32_t> pairs(2 * 12);
        int32_t used = 0;
        const int np = orc_rank_pairs(12, 9, R.data(), 0.25f, b & 1, u.data(), (int32_t)u.size(), pairs.data(), &used);
        CHECK(np >= 0 && np <= 9);
    }
}

void configuration() {
    // tut_SexualPop from module attributes, then each rule's error case
    const char *fk[] = {"Fertility_min_age", "Fertility_max_age", "Fertility_interbirth"};
    const char *fv[] = {"15", "45", "3"};
    const char *vk[] = {"Verhulst_b0", "Verhulst_d0", "Verhulst_theta", "Verhulst_K"};
    const char *vv[] = {"0.8", "0.001", "0.1", "30"};
    const char *rk[] = {"RandomMove_prob"};
    const char *rv[] = {"0.1"};
    const char *ak[] = {"ATanDeath_max_age", "ATanDeath_range", "ATanDeath_slope"};
    const char *av[] = {"60", "10", "0.2"};
    qhg_module_attrs mods[] = {{"Fertility", 3, fk, fv}, {"Verhulst", 4, vk, vv}, {"RandomMove", 1, rk, rv},
                               {"ATanDeath", 3, ak, av}, {"GetOld", 0, nullptr, nullptr},
                               {"RandomPair", 0, nullptr, nullptr}};
    const char *pn[] = {"GetOld", "Fertility", "RandomPair", "Verhulst", "RandomMove"};
    const int32_t pv[] = {1, 2, 3, 4, 5};
    qhg_params P;
    CHECK(qhg_params_from_modules("tut_SexualPop", mods, 6, pn, pv, 5, 3, &P) == 0);
    double v = 0;
    CHECK(qhg_params_get_attribute(&P, "Verhulst_b0", &v) == 0 && v == 0.8);
    CHECK(qhg_params_set_attribute(&P, "Verhulst_b0", 0.5) == 0);
    CHECK(qhg_params_get_attribute(&P, "No_such", &v) != 0);
    const char *bad[] = {"x", "45", "3"};
    qhg_module_attrs m2[6];
    std::memcpy(m2, mods, sizeof mods);
    m2[0] = qhg_module_attrs{"Fertility", 2, fk, fv};                  // missing attribute
    CHECK(qhg_params_from_modules("tut_SexualPop", m2, 6, pn, pv, 5, 3, &P) != 0);
    m2[0] = qhg_module_attrs{"Fertility", 3, fk, bad};                 // not a number
    CHECK(qhg_params_from_modules("tut_SexualPop", m2, 6, pn, pv, 5, 3, &P) != 0);
    CHECK(qhg_params_from_modules("NoSuchPop", mods, 6, pn, pv, 5, 3, &P) != 0);
    const char *pn2[] = {"Navigate"};
    CHECK(qhg_params_from_modules("tut_SexualPop", mods, 6, pn2, pv, 1, 3, &P) != 0);
    for (const char *cls : {"tut_SexualPop", "tut_EnvironCapAltPop", "OoAPop", "OoANavGen2bitPop",
                            "OoANavGen2bitAssHamPop", "OoANavGen2bitAssPop", "NoSuchPop"}) {
        int32_t kind = 0;
        for (int i = 0; qhg_params_class_action(cls, i, &kind); i++) {}
        for (int e = 0; e < 4; e++) {
            (void)qhg_params_evaluator_name(cls, e);
            (void)qhg_params_weight_name(cls, e);
            (void)qhg_params_polyline_name(cls, e);
        }
    }
    for (int kind = QHG_ACT_GETOLD; kind <= QHG_ACT_LAST; kind++)
        for (int i = 0; qhg_params_attribute_name(kind, i); i++) {}
    std::printf("configuration ok\n");
}

void qdf(const char *dir) {
    const World w = make_world(10, 8, 3);
    qhg_params P = base_params(21);
    add_action(P, QHG_ACT_GETOLD, 1, "GetOld");
    add_action(P, QHG_ACT_FERTILITY, 2, "Fertility");
    add_action(P, QHG_ACT_RANDOMPAIR, 3, "RandomPair");
    add_action(P, QHG_ACT_VERHULST, 4, "Verhulst");
    add_action(P, QHG_ACT_RANDOMMOVE, 5, "RandomMove");
    add_action(P, QHG_ACT_ATANDEATH, 6, "ATanDeath");
    Pop a = make_agents(500, w.nc, 0, 4);
    qhg_agents v = a.view(0);
    const std::string f = std::string(dir) + "/san.qdf";
    CHECK(qhg_qdf_write_agents(f.c_str(), 1, "sapiens", "tut_SexualPop", &P, w.nc, nullptr, 12, 12.0, 500, &v, 0) == 0);
    qhg_qdf_info info{};
    CHECK(qhg_qdf_read_info(f.c_str(), "sapiens", &info) == 0 && info.num_agents == 500 && info.has_params);
    qhg_params Q;
    CHECK(qhg_qdf_read_params(f.c_str(), "sapiens", 0, &Q) == 0);
    // the file holds the class's actions and their attributes
    CHECK(Q.num_actions == P.num_actions && Q.seed == P.seed);
    for (int i = 0; i < P.num_actions; i++) {
        int j = 0;
        while (j < Q.num_actions && Q.action_kind[j] != P.action_kind[i]) j++;
        CHECK(j < Q.num_actions && Q.action_prio[j] == P.action_prio[i]);
        for (int k = 0; const char *name = qhg_params_attribute_name(P.action_kind[i], k); k++) {
            double x = 0, y = 0;
            CHECK(qhg_params_get_attribute(&P, name, &x) == 0 && qhg_params_get_attribute(&Q, name, &y) == 0);
            CHECK(std::memcmp(&x, &y, sizeof x) == 0);
        }
    }
    Pop b = make_agents(500, w.nc, 0, 99);
    qhg_agents bv = b.view(0);
    CHECK(qhg_qdf_read_agents(f.c_str(), "sapiens", &bv) == 0);
    CHECK(b.id == a.id && b.cell == a.cell && b.birth == a.birth);
    // an empty population (an extinct one): no arrays needed either way
    qhg_agents none{0, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    const std::string fe = std::string(dir) + "/empty.qdf";
    CHECK(qhg_qdf_write_agents(fe.c_str(), 1, "sapiens", "tut_SexualPop", &P, w.nc, nullptr, 40, 40.0, 700, &none,
                               0) == 0);
    CHECK(qhg_qdf_read_info(fe.c_str(), "sapiens", &info) == 0 && info.num_agents == 0 && info.max_id == 700);
    CHECK(qhg_qdf_read_agents(fe.c_str(), "sapiens", &none) == 0);
    CHECK(qhg_qdf_read_info(f.c_str(), "nobody", &info) != 0);
    CHECK(qhg_qdf_read_info((std::string(dir) + "/missing.qdf").c_str(), "sapiens", &info) != 0);
    qhg_qdf_grid_info gi{};
    CHECK(qhg_qdf_read_grid_info(f.c_str(), &gi) != 0);   // no grid group in a population file
    std::printf("qdf ok\n");
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 2) {
        std::fprintf(stderr, "usage: sanitize_driver <scratch dir>\n");
        return 2;
    }
    oracle_units();
    oracle_models();
    configuration();
    qdf(argv[1]);
    if (failures) {
        std::fprintf(stderr, "%d checks failed\n", failures);
        return 1;
    }
    std::printf("all checks passed\n");
    return 0;
}
