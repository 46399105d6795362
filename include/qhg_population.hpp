/*
 * qhg_population.hpp -- C++ host adapter over the C ABI (header only).
 *
 * qhg::Population owns one device population and offers the calls the
 * reference's driver makes on a population, with the reference's method
 * names, argument meaning and error convention (int, 0 = OK, negative =
 * error; lastError() gives the text):
 *
 *   PopBase pure virtuals            QHG4/core/PopBase.h:16-121
 *   PopLooper::doStep call order     QHG4/core/PopLooper.cpp:166-202
 *   Simulator::runLoop counts/events QHG4/app/Simulator.cpp:296-374
 *   SPopulation::readSpeciesData     QHG4/core/SPopulation.cpp:1108-1146
 *   SPopulation::addAgent(Data)      QHG4/core/SPopulation.cpp:1149-1222
 *
 * It needs nothing but qhg_gpu.h and libqhg4_amd.so.  The PopBase plugin
 * (qhg4_amd/plugin) forwards the reference's virtuals to it.
 *
 * Lifecycle: construct with the grid and environment (borrowed: the caller
 * keeps the arrays alive; flushEvents reads them again, as the reference's
 * actions read the grid's Geography, Climate and Vegetation in place), then
 * readSpeciesData (or setParams) creates the device population, addAgent
 * stages agents, preLoop uploads them, and the step calls follow.
 */
#ifndef QHG_POPULATION_HPP
#define QHG_POPULATION_HPP

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <set>
#include <string>
#include <utility>
#include <vector>

#include "qhg_gpu.h"

namespace qhg {

// event ids of QHG4/utils_qhg/EventConsts.h
constexpr int kEventIdGeo = 2;
constexpr int kEventIdClimate = 3;
constexpr int kEventIdVeg = 4;
constexpr int kEventIdNav = 5;
constexpr int kEventIdFlush = 20;

// one module's attributes (ParamProvider2's ModuleComplex::getAttributes)
using AttrMap = std::map<std::string, std::string>;
// modules by name, sub-modules listed beside their parents
// ("MultiEvaluator[Alt+NPP]", "SingleEvaluator[Alt]", ...)
using ModuleList = std::vector<std::pair<std::string, AttrMap>>;

// The parameter block of a species class from its modules and priorities
// (SPopulation::readSpeciesData -> getPrioInfos + Prioritizer::
// getActionAttributes).  Priorities are the XML's strings.
inline int params_from_modules(const std::string &pop_class, const ModuleList &mods, const AttrMap &prios,
                               uint64_t seed, qhg_params *out) {
    std::vector<std::vector<const char *>> keys(mods.size()), vals(mods.size());
    std::vector<qhg_module_attrs> m(mods.size());
    for (size_t i = 0; i < mods.size(); i++) {
        for (const auto &kv : mods[i].second) {
            keys[i].push_back(kv.first.c_str());
            vals[i].push_back(kv.second.c_str());
        }
        m[i] = qhg_module_attrs{mods[i].first.c_str(), (int32_t)keys[i].size(), keys[i].data(), vals[i].data()};
    }
    std::vector<const char *> pn;
    std::vector<int32_t> pv;
    for (const auto &kv : prios) {
        char *end = nullptr;
        const long v = std::strtol(kv.second.c_str(), &end, 10);
        if (end == kv.second.c_str()) return QHG_ERR_ARG;
        pn.push_back(kv.first.c_str());
        pv.push_back((int32_t)v);
    }
    return qhg_params_from_modules(pop_class.c_str(), m.data(), (int32_t)m.size(), pn.data(), pv.data(),
                                   (int32_t)pn.size(), seed, out);
}

// Agents as arrays (the download / upload layout of the C ABI).
struct Agents {
    std::vector<int64_t> id, mate_id;
    std::vector<int32_t> cell, num_babies;
    std::vector<uint8_t> life, gender;
    std::vector<double> birth_time, age, last_birth;
    std::vector<uint64_t> genome;   // size() * genome_words
    int32_t genome_words = 0;

    size_t size() const { return id.size(); }
    void resize(size_t n, int32_t gw) {
        id.resize(n);
        mate_id.resize(n, -3);
        cell.resize(n);
        num_babies.resize(n);
        life.resize(n);
        gender.resize(n);
        birth_time.resize(n);
        age.resize(n);
        last_birth.resize(n);
        genome_words = gw;
        genome.resize(n * (size_t)gw);
    }
    qhg_agents view() {
        return qhg_agents{(int64_t)size(), id.data(),         cell.data(),       life.data(),
                          gender.data(),   birth_time.data(), age.data(),        last_birth.data(),
                          mate_id.data(),  num_babies.data(), genome_words ? genome.data() : nullptr};
    }
};

// One agent line of the reference's agent data ("LifeState AgentID
// BirthTime Gender" then the class's fields Age and LastBirth, separators
// " ,;:"; SPopulation::addAgentData and e.g. OoAPop::addPopSpecificAgentData,
// QHG4/populations/OoAPop.cpp:245-256).  Returns 0 or QHG_ERR_ARG.
inline int parse_agent_data(const char *data, double out[6], int *count) {
    int k = 0;
    const char *p = data;
    while (p && *p && k < 6) {
        while (*p && std::strchr(" ,;:\t\n", *p)) p++;
        if (!*p) break;
        char *end = nullptr;
        out[k] = std::strtod(p, &end);
        if (end == p || (*end && !std::strchr(" ,;:\t\n", *end))) return QHG_ERR_ARG;
        k++;
        p = end;
    }
    *count = k;
    return k >= 4 ? 0 : QHG_ERR_ARG;
}

class Population {
public:
    Population(const qhg_grid_desc &grid, const qhg_env_desc &env, std::string pop_class, std::string species)
        : grid_(grid), env_(env), class_(std::move(pop_class)), species_(std::move(species)) {}
    ~Population() {
        if (pop_) qhg_gpu_destroy(pop_);
    }
    Population(const Population &) = delete;
    Population &operator=(const Population &) = delete;

    // -- species data ------------------------------------------------------
    int readSpeciesData(const ModuleList &mods, const AttrMap &prios, uint64_t seed) {
        qhg_params p;
        const int rc = params_from_modules(class_, mods, prios, seed, &p);
        if (rc) return fail(rc, "readSpeciesData");
        return setParams(p);
    }
    // a ready parameter block (e.g. read from a QDF species group)
    int setParams(const qhg_params &p) {
        if (pop_) return fail(QHG_ERR_STATE, "setParams: the population exists already");
        params_ = p;
        const int rc = qhg_gpu_create(&grid_, &env_, &params_, capacity_, &pop_);
        if (rc) return fail(rc, "qhg_gpu_create");
        for (int i = 0; i < params_.num_actions; i++)
            if (params_.action_kind[i] == QHG_ACT_GENETICS) gwords_ = 2 * qhg_gpu_genome_words(params_.gen_genome_size);
        return 0;
    }
    const std::string &getClassName() const { return class_; }
    const std::string &getSpeciesName() const { return species_; }
    const qhg_params &params() const { return params_; }
    qhg_pop *handle() const { return pop_; }
    int32_t genomeWords() const { return gwords_; }

    // -- agents ------------------------------------------------------------
    int reserveAgentSpace(int64_t n) {
        if (n > capacity_) capacity_ = n;
        return 0;
    }
    // one agent line (see parse_agent_data) into a cell; staged until the
    // next preLoop, or uploaded at once after it
    int addAgent(int cell, const char *data) {
        double v[6];
        int k = 0;
        if (parse_agent_data(data, v, &k)) return fail(QHG_ERR_ARG, std::string("addAgent: cannot read ") + data);
        const size_t i = staged_.size();
        staged_.resize(i + 1, gwords_);
        staged_.cell[i] = cell;
        staged_.life[i] = (uint8_t)v[0];
        staged_.id[i] = (int64_t)v[1];
        staged_.birth_time[i] = v[2];
        staged_.gender[i] = (uint8_t)v[3];
        staged_.age[i] = k > 4 ? v[4] : 0.0;
        staged_.last_birth[i] = k > 5 ? v[5] : 0.0;
        return looping_ ? flushStaged() : 0;
    }
    // agents given as arrays (genomes included when the class has Genetics)
    int addAgents(const Agents &a) {
        if (a.genome_words != gwords_) return fail(QHG_ERR_ARG, "addAgents: genome words differ");
        const size_t i0 = staged_.size(), n = a.size();
        staged_.resize(i0 + n, gwords_);
        for (size_t k = 0; k < n; k++) {
            staged_.id[i0 + k] = a.id[k];
            staged_.cell[i0 + k] = a.cell[k];
            staged_.life[i0 + k] = a.life[k];
            staged_.gender[i0 + k] = a.gender[k];
            staged_.birth_time[i0 + k] = a.birth_time[k];
            staged_.age[i0 + k] = a.age[k];
            staged_.last_birth[i0 + k] = a.last_birth[k];
            staged_.num_babies[i0 + k] = a.num_babies.empty() ? 0 : a.num_babies[k];
        }
        if (gwords_) std::memcpy(&staged_.genome[i0 * gwords_], a.genome.data(), n * gwords_ * 8);
        return looping_ ? flushStaged() : 0;
    }
    // the live agents, id order
    int download(Agents &out) {
        int64_t n = 0;
        if (int rc = need()) return rc;
        if (int rc = qhg_gpu_num_alive(pop_, &n)) return fail(rc, "qhg_gpu_num_alive");
        out.resize((size_t)n, gwords_);
        qhg_agents v = out.view();
        if (int rc = qhg_gpu_download_agents(pop_, &v)) return fail(rc, "qhg_gpu_download_agents");
        return 0;
    }

    // -- the loop ------------------------------------------------------------
    int setPrioList() { return need(); }
    // the priority levels of the scheduled actions; returns their number
    unsigned getPrios(std::set<unsigned> &prios) const {
        prios.clear();
        for (int i = 0; i < params_.num_actions; i++)
            if (params_.action_prio[i] >= 0) prios.insert((unsigned)params_.action_prio[i]);
        return (unsigned)prios.size();
    }
    int preLoop() {
        if (int rc = need()) return rc;
        if (int rc = flushStaged()) return rc;
        if (int rc = qhg_gpu_pre_loop(pop_, 0.0)) return fail(rc, "qhg_gpu_pre_loop");
        looping_ = true;
        return refresh();
    }
    int postLoop() { return sync(); }
    int initializeStep(float t) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_initialize_step(pop_, (double)t);
        return rc ? fail(rc, "qhg_gpu_initialize_step") : 0;
    }
    int doActions(unsigned prio, float t) {
        const int rc = qhg_gpu_do_actions(pop_, prio, (double)t);
        return rc ? fail(rc, "qhg_gpu_do_actions") : 0;
    }
    int finalizeStep() {
        const int rc = qhg_gpu_finalize_step(pop_);
        counts_fresh_ = false;
        return rc ? fail(rc, "qhg_gpu_finalize_step") : 0;
    }
    int preWrite(float) { return refresh(); }
    // the clock of a resumed run (SPopulation::restoreSpeciesDataQDF,
    // QHG4/core/SPopulation.cpp:2138-2207): the next step's index and time
    // and the highest id given so far
    int setClock(int64_t step, double t, int64_t max_id) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_set_clock(pop_, step, t, max_id);
        counts_fresh_ = false;
        return rc ? fail(rc, "qhg_gpu_set_clock") : 0;
    }
    // whole steps without host round trips (the device keeps the clock)
    int run(int nsteps) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_run(pop_, nsteps);
        counts_fresh_ = false;
        return rc ? fail(rc, "qhg_gpu_run") : 0;
    }

    // -- counts ------------------------------------------------------------
    unsigned long getNumAgents(int cell) {
        if (!counts_fresh_ && refresh()) return 0;
        return (cell >= 0 && cell < grid_.num_cells) ? counts_[(size_t)cell] : 0;
    }
    unsigned long getNumAgentsTotal() {
        if (!counts_fresh_ && refresh()) return 0;
        return (unsigned long)total_;
    }
    unsigned long getNumAgentsEffective() { return getNumAgentsTotal(); }
    unsigned long getNumAgentsMax() { return getNumAgentsTotal(); }   // no holes in the buffers
    void updateTotal() { refresh(); }
    void updateNumAgentsPerCell() { refresh(); }
    int getNumCells() const { return grid_.num_cells; }
    long getMaxLoadedID() {
        int64_t step = 0, max_id = 0;
        double t = 0;
        if (!pop_ || qhg_gpu_get_clock(pop_, &step, &t, &max_id)) return -1;
        return (long)max_id;
    }
    int stats(qhg_stats *s) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_stats(pop_, s);
        return rc ? fail(rc, "qhg_gpu_stats") : 0;
    }

    // -- attributes, events, housekeeping ----------------------------------
    int modifyAttributes(const std::string &name, double value) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_modify_attribute(pop_, name.c_str(), value);
        if (rc) return fail(rc, "modifyAttributes " + name);
        qhg_params_set_attribute(&params_, name.c_str(), value);
        return 0;
    }
    // EVENT_ID_GEO / CLIMATE / VEG mark the environment changed; the
    // population reacts at flushEvents.  EVENT_ID_NAV marks the navigation
    // changed: the caller passes the new one with setNavigation
    // (takeNavEvent tells it so).  Other ids are not environment events.
    int updateEvent(int event_id, float) {
        if (event_id == kEventIdGeo) pending_ |= QHG_EVENT_GEO;
        if (event_id == kEventIdClimate) pending_ |= QHG_EVENT_CLIMATE;
        if (event_id == kEventIdVeg) pending_ |= QHG_EVENT_VEG;
        if (event_id == kEventIdNav) nav_event_ = true;
        return 0;
    }
    bool takeNavEvent() {
        const bool e = nav_event_;
        nav_event_ = false;
        return e;
    }
    // the grid's sea-ways and bridges (Navigation::setData / setBridges), for
    // Navigate: before preLoop the grid's data, afterwards a navigation
    // event with its flush (the arrays are copied)
    int setNavigation(const qhg_nav_desc &nav) {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_set_navigation(pop_, &nav);
        return rc ? fail(rc, "qhg_gpu_set_navigation") : 0;
    }
    int flushEvents(float) {
        if (!pending_ || !pop_) return 0;
        const int flags = qhg_params_event_flags(class_.c_str(), pending_);
        pending_ = 0;
        if (flags < 0) return fail(flags, "qhg_params_event_flags");
        const int rc = qhg_gpu_env_event(pop_, &env_, flags);
        counts_fresh_ = false;
        return rc ? fail(rc, "qhg_gpu_env_event") : 0;
    }
    // the environment arrays moved (a new Geography etc.): later events read these
    void setEnvironment(const qhg_env_desc &env) { env_ = env; }
    int flushDeadSpace() { return 0; }   // deaths are committed inside each step
    void compactData() {
        if (pop_) qhg_gpu_compact(pop_);
    }
    int sync() {
        if (int rc = need()) return rc;
        const int rc = qhg_gpu_sync(pop_);
        return rc ? fail(rc, "qhg_gpu_sync") : 0;
    }
    const std::string &lastError() const { return err_; }

private:
    int need() {
        return pop_ ? 0 : fail(QHG_ERR_STATE, "no device population (readSpeciesData first)");
    }
    int fail(int rc, const std::string &what) {
        const char *e = qhg_gpu_last_error();
        err_ = what + ": " + (e ? e : "");
        return rc;
    }
    int refresh() {
        if (int rc = need()) return rc;
        counts_.assign((size_t)grid_.num_cells, 0u);
        int64_t n = 0;
        if (int rc = qhg_gpu_cell_counts(pop_, counts_.data())) return fail(rc, "qhg_gpu_cell_counts");
        if (int rc = qhg_gpu_num_alive(pop_, &n)) return fail(rc, "qhg_gpu_num_alive");
        total_ = n;
        counts_fresh_ = true;
        return 0;
    }
    // staged agents join the device population (the whole population is
    // uploaded again when it already holds agents; the clock stays)
    int flushStaged() {
        if (staged_.size() == 0) return 0;
        if (int rc = need()) return rc;
        int64_t step = 0, max_id = 0, n = 0;
        double t = 0;
        if (int rc = qhg_gpu_get_clock(pop_, &step, &t, &max_id)) return fail(rc, "qhg_gpu_get_clock");
        if (int rc = qhg_gpu_num_alive(pop_, &n)) return fail(rc, "qhg_gpu_num_alive");
        Agents all;
        if (n > 0) {
            if (int rc = download(all)) return rc;
        }
        const size_t i0 = all.size();
        all.resize(i0 + staged_.size(), gwords_);
        for (size_t k = 0; k < staged_.size(); k++) {
            all.id[i0 + k] = staged_.id[k];
            all.cell[i0 + k] = staged_.cell[k];
            all.life[i0 + k] = staged_.life[k];
            all.gender[i0 + k] = staged_.gender[k];
            all.birth_time[i0 + k] = staged_.birth_time[k];
            all.age[i0 + k] = staged_.age[k];
            all.last_birth[i0 + k] = staged_.last_birth[k];
            all.num_babies[i0 + k] = staged_.num_babies[k];
            all.mate_id[i0 + k] = -3;
        }
        if (gwords_) std::memcpy(&all.genome[i0 * gwords_], staged_.genome.data(), staged_.genome.size() * 8);
        qhg_agents v = all.view();
        if (int rc = qhg_gpu_upload_agents(pop_, &v)) return fail(rc, "qhg_gpu_upload_agents");
        if (n > 0 || step > 0) {
            if (int rc = qhg_gpu_set_clock(pop_, step, t, max_id)) return fail(rc, "qhg_gpu_set_clock");
        }
        staged_ = Agents{};
        counts_fresh_ = false;
        return 0;
    }

    qhg_grid_desc grid_;
    qhg_env_desc env_;
    std::string class_, species_;
    qhg_params params_{};
    qhg_pop *pop_ = nullptr;
    int64_t capacity_ = 0;
    int32_t gwords_ = 0;
    Agents staged_;
    bool looping_ = false;
    int pending_ = 0;
    bool nav_event_ = false;
    std::vector<uint32_t> counts_;
    int64_t total_ = 0;
    bool counts_fresh_ = false;
    std::string err_;
};

}  // namespace qhg

#endif
